// Integrator kernels for gfx950 (MI355X).
//
// One particle per lane in a persistent kernel: every lane integrates a
// particle until it reaches the end of the segment, hands the state back to
// HBM and pulls the next particle index from a device work queue (one
// wave-aggregated atomic per refill), so particles whose step counts differ
// by decades share waves without idling lanes.  All stage arithmetic of an
// attempt is the same instruction stream for accepted and rejected lanes;
// only bookkeeping diverges.
//
// The arithmetic restates the reference per particle (gkowal/particles,
// sources/integrations.F90, accelerations.F90, interpolations.F90) in its
// operation order; this file is compiled with -ffp-contract=off so every
// FP64 add/mul/div/sqrt rounds like the reference's own build, and the
// step-size powers use ptc_libm_pow (libm_pow.hpp), which reproduces the
// host libm's pow bit for bit.  Trajectories and step counters then match
// the reference bit for bit.
//
// Redundant evaluations of the reference are not repeated, which leaves every
// result unchanged:
//   * RK5 recomputes k6 = dt f(s6) on accept (integrations.F90:156) -- the
//     stored k6 is identical;
//   * DOP853 recomputes f(si) (:445, equal to stage 1) and f(sf) (:446, equal
//     to the next step's stage 1) -- f(sf) is evaluated once and reused;
//   * dti_guess evaluates f at a trial point and never uses it (:576);
//   * a rejected attempt reuses f(si) (stage 1 does not depend on dt).
#include "trio_impl.hpp"

namespace ptc {
namespace {

__global__ void totals_kernel(const DeviceParticles S, unsigned long long *out)
{
    const int64_t N = S.np;
    unsigned long long a = 0, b = 0, c = 0, d = 0, e = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * blockDim.x) {
        a += (unsigned long long)S.cnt[i];
        b += (unsigned long long)S.cnt[N + i];
        c += (unsigned long long)S.cnt[2 * N + i];
        d += (unsigned long long)S.nrhs[i];
        e += S.status[i] == PTC_STATUS_RUNNING ? 1ull : 0ull;
    }
    // wave reduction, then one atomic per wave and counter
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off);
        b += __shfl_down(b, off);
        c += __shfl_down(c, off);
        d += __shfl_down(d, off);
        e += __shfl_down(e, off);
    }
    if (__lane_id() == 0) {
        atomicAdd(out + 0, a);
        atomicAdd(out + 1, b);
        atomicAdd(out + 2, c);
        atomicAdd(out + 3, d);
        atomicAdd(out + 4, e);
    }
}

__global__ void zero_i32_kernel(int32_t *p, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

__global__ void zero_i64_kernel(int64_t *p, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

__global__ void clear_flag_kernel(int32_t *p, int64_t n, int32_t bit)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] &= ~bit;
}

__global__ void to_aos_kernel(const double *__restrict__ soa, int64_t nodes, double *__restrict__ aos)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nodes;
         i += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 6; c++) aos[i * 6 + c] = soa[c * nodes + i];
    }
}

// one thread per value of the brick array: cell-major, then corner, then
// component, so the writes are contiguous
__global__ void to_bricks_kernel(const double *__restrict__ aos, int dm0, int dm1, int dm2,
                                 int64_t nodes, double *__restrict__ bricks)
{
    const int64_t total = nodes * 48;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < total;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t cell = g / 48;
        const int rem = (int)(g - cell * 48), j = rem / 6, comp = rem - j * 6;
        const int i = (int)(cell % dm0);
        const int jy = (int)((cell / dm0) % dm1);
        const int kz = (int)(cell / ((int64_t)dm0 * dm1));
        // corner j = (x, y, z) offsets (j>>2, (j>>1)&1, j&1): x slowest
        const int ii = (j >> 2) ? (i + 1 == dm0 ? 0 : i + 1) : i;
        const int jj = ((j >> 1) & 1) ? (jy + 1 == dm1 ? 0 : jy + 1) : jy;
        const int kk = (j & 1) ? (kz + 1 == dm2 ? 0 : kz + 1) : kz;
        const int64_t node = ii + (int64_t)dm0 * (jj + (int64_t)dm1 * kk);
        bricks[g] = aos[node * 6 + comp];
    }
}

// SOA planes (FP64) -> FP32 nodes (U or E, B, |B| from the FP64 values, 0)
__global__ void to_nodes_f32_kernel(const double *__restrict__ soa, int64_t nodes,
                                    float *__restrict__ out)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nodes;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double bx = soa[3 * nodes + i], by = soa[4 * nodes + i], bz = soa[5 * nodes + i];
        float4 a, b;
        a.x = float(soa[i]);
        a.y = float(soa[nodes + i]);
        a.z = float(soa[2 * nodes + i]);
        a.w = float(bx);
        b.x = float(by);
        b.y = float(bz);
        b.z = float(sqrt(bx * bx + by * by + bz * bz));
        b.w = 0.0f;
        ((float4 *)out)[2 * i] = a;
        ((float4 *)out)[2 * i + 1] = b;
    }
}

// SOA planes (FP64) -> FP32 pair rows (64 bytes): value c of node n at
// position 2 c, of its +x neighbour (periodic in x) at 2 c + 1, then four
// zeros
__global__ void to_pairs_f32_kernel(const double *__restrict__ soa, int dm0, int64_t nodes,
                                    float *__restrict__ out)
{
    for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < nodes;
         n += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(n % dm0);
        const int64_t nx = n - i + (i + 1 == dm0 ? 0 : i + 1);
        float *__restrict__ row = out + 16 * n;
#pragma unroll
        for (int c = 0; c < 6; c++) {
            row[2 * c] = float(soa[c * nodes + n]);
            row[2 * c + 1] = float(soa[c * nodes + nx]);
        }
#pragma unroll
        for (int c = 12; c < 16; c++) row[c] = 0.0f;
    }
}

// SOA planes (FP64) -> FP64 pair rows (128 bytes): value c of node n at
// position 2 c, of its +x neighbour (periodic in x) at 2 c + 1, then four
// zeros
__global__ void to_pairs_f64_kernel(const double *__restrict__ soa, int dm0, int64_t nodes,
                                    double *__restrict__ out)
{
    for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < nodes;
         n += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(n % dm0);
        const int64_t nx = n - i + (i + 1 == dm0 ? 0 : i + 1);
        double *__restrict__ row = out + 16 * n;
#pragma unroll
        for (int c = 0; c < 6; c++) {
            row[2 * c] = soa[c * nodes + n];
            row[2 * c + 1] = soa[c * nodes + nx];
        }
#pragma unroll
        for (int c = 12; c < 16; c++) row[c] = 0.0;
    }
}

// SOA planes (FP64) -> FP32 quad rows (128 bytes): value c of node a of
// n = (i, j, k), (i+1, j, k), (i, j+1, k), (i+1, j+1, k) (periodic), a = 0..3,
// at position 4 c + a, then eight zeros
__global__ void to_quads_f32_kernel(const double *__restrict__ soa, int dm0, int dm1,
                                    int64_t nodes, float *__restrict__ out)
{
    for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < nodes;
         n += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(n % dm0);
        const int j = (int)((n / dm0) % dm1);
        const int64_t base = n - i - (int64_t)j * dm0;  // node (0, 0, k)
        const int i1 = i + 1 == dm0 ? 0 : i + 1, j1 = j + 1 == dm1 ? 0 : j + 1;
        const int64_t q[4] = {n, base + i1 + (int64_t)j * dm0, base + i + (int64_t)j1 * dm0,
                              base + i1 + (int64_t)j1 * dm0};
        float *__restrict__ row = out + 32 * n;
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int c = 0; c < 6; c++) row[4 * c + a] = float(soa[c * nodes + q[a]]);
#pragma unroll
        for (int c = 24; c < 32; c++) row[c] = 0.0f;
    }
}

// SOA planes (FP64) -> FP64 column rows (192 bytes, stored z fastest):
// value c of node a of n = (i, j, k), (i+1, j, k), (i, j+1, k), (i+1, j+1, k)
// (periodic), a = 0..3, at position 4 c + a of row (i + dm0 j) dm2 + k.
// One thread per row, rows in storage order.
__global__ void to_columns_f64_kernel(const double *__restrict__ soa, int dm0, int dm1, int dm2,
                                      int64_t nodes, double *__restrict__ out)
{
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nodes;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t ij = r / dm2;
        const int k = (int)(r - ij * dm2);
        const int i = (int)(ij % dm0), j = (int)(ij / dm0);
        const int64_t base = (int64_t)k * dm0 * dm1;  // node (0, 0, k)
        const int i1 = i + 1 == dm0 ? 0 : i + 1, j1 = j + 1 == dm1 ? 0 : j + 1;
        const int64_t q[4] = {base + i + (int64_t)j * dm0, base + i1 + (int64_t)j * dm0,
                              base + i + (int64_t)j1 * dm0, base + i1 + (int64_t)j1 * dm0};
        double *__restrict__ row = out + 24 * r;
#pragma unroll
        for (int c = 0; c < 6; c++)
#pragma unroll
            for (int a = 0; a < 4; a++) row[4 * c + a] = soa[c * nodes + q[a]];
    }
}

// SOA planes (FP64) -> FP32 column rows (128 bytes, stored z fastest): the
// FP32 quad row of node (i, j, k) (value c of node a at 4 c + a, then eight
// zeros) at row (i + dm0 j) dm2 + k.  One thread per row, in storage order.
__global__ void to_columns_f32_kernel(const double *__restrict__ soa, int dm0, int dm1, int dm2,
                                      int64_t nodes, float *__restrict__ out)
{
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nodes;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t ij = r / dm2;
        const int k = (int)(r - ij * dm2);
        const int i = (int)(ij % dm0), j = (int)(ij / dm0);
        const int64_t base = (int64_t)k * dm0 * dm1;  // node (0, 0, k)
        const int i1 = i + 1 == dm0 ? 0 : i + 1, j1 = j + 1 == dm1 ? 0 : j + 1;
        const int64_t q[4] = {base + i + (int64_t)j * dm0, base + i1 + (int64_t)j * dm0,
                              base + i + (int64_t)j1 * dm0, base + i1 + (int64_t)j1 * dm0};
        float *__restrict__ row = out + 32 * r;
#pragma unroll
        for (int c = 0; c < 6; c++)
#pragma unroll
            for (int a = 0; a < 4; a++) row[4 * c + a] = float(soa[c * nodes + q[a]]);
#pragma unroll
        for (int c = 24; c < 32; c++) row[c] = 0.0f;
    }
}

// value comp (0..5) of node n (x fastest) of a stored static field, any layout
__device__ __forceinline__ double stored_value(const DeviceFields &F, int64_t n, int comp)
{
    const int64_t rv = row_values(F.layout, F.prec);
    const int pos = row_nodes(F.layout) * comp;
    const int64_t r = row_of(F.layout, F.dm, n);
    if (F.prec == PTC_FP32) return double(((const float *)F.data)[r * rv + pos]);
    if (F.layout == LAYOUT_SOA) return ((const double *)F.data)[comp * F.nodes + n];
    return ((const double *)F.data)[r * rv + pos];
}

// Rows of a new static field (layout, prec) from a stored one (any layout
// and precision), node by node: node rows (FP64: six values; FP32: six, |B|
// from the stored values in FP64, 0), pair rows (the node, its +x
// neighbour, zeros), FP32 quad rows (the node, +x, +y, +x+y, zeros; value
// c of the row's node a at row_nodes * c + a) or column rows (quad rows in
// row row_of(n); FP64 without zeros), the neighbours periodic.  The
// same values as building the new form from the stored field's planes.
__global__ void convert_rows_kernel(const DeviceFields F, void *__restrict__ out, int layout,
                                    int prec)
{
    const int dm0 = F.dm[0], dm1 = F.dm[1];
    const int64_t nodes = F.nodes;
    const int rv = row_values(layout, prec);
    for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < nodes;
         n += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(n % dm0);
        const int j = (int)((n / dm0) % dm1);
        const int64_t base = n - i - (int64_t)j * dm0;
        const int i1 = i + 1 == dm0 ? 0 : i + 1, j1 = j + 1 == dm1 ? 0 : j + 1;
        const int64_t q[4] = {n, base + i1 + (int64_t)j * dm0, base + i + (int64_t)j1 * dm0,
                              base + i1 + (int64_t)j1 * dm0};
        const int nq = row_nodes(layout);
        const int64_t ro = row_of(layout, F.dm, n);  // the new row of node n
        for (int a = 0; a < 4; a++) {
            if (a >= nq) break;
#pragma unroll
            for (int c = 0; c < 6; c++) {
                const double v = stored_value(F, q[a], c);
                if (prec == PTC_FP32) ((float *)out)[ro * rv + nq * c + a] = float(v);
                else ((double *)out)[ro * rv + nq * c + a] = v;
            }
        }
        if (prec == PTC_FP32 && layout == LAYOUT_AOS) {
            const double bx = stored_value(F, n, 3), by = stored_value(F, n, 4),
                         bz = stored_value(F, n, 5);
            ((float *)out)[n * rv + 6] = float(sqrt(bx * bx + by * by + bz * bz));
            ((float *)out)[n * rv + 7] = 0.0f;
        } else {
            for (int c = 6 * nq; c < rv; c++) {
                if (prec == PTC_FP32) ((float *)out)[ro * rv + c] = 0.0f;
                else ((double *)out)[ro * rv + c] = 0.0;
            }
        }
    }
}

// One component of a stored field back to a plane of doubles (the
// reference's uu(:,:,:,c) / bb(:,:,:,c) order): from the six SOA planes,
// FP64 node rows, FP32 node rows (eight floats), FP32 pair rows (sixteen
// floats), FP32 quad rows (32 floats), FP64 pair rows (16 doubles) or
// column rows (row row_of(i): 32 floats or 24 doubles); in pair, quad and
// column rows the row's own node's value comp at row_nodes * comp.
__global__ void extract_plane_kernel(const void *__restrict__ data, int layout, int prec, int dm0,
                                     int dm1, int dm2, int64_t nodes, int comp,
                                     double *__restrict__ out)
{
    const int64_t rv = row_values(layout, prec);
    const int pos = row_nodes(layout) * comp;
    const int dm[3] = {dm0, dm1, dm2};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nodes;
         i += (int64_t)gridDim.x * blockDim.x) {
        double v;
        const int64_t r = row_of(layout, dm, i);
        if (prec == PTC_FP32)
            v = double(((const float *)data)[r * rv + pos]);
        else if (layout == LAYOUT_SOA)
            v = ((const double *)data)[comp * nodes + i];
        else
            v = ((const double *)data)[r * rv + pos];
        out[i] = v;
    }
}

}  // namespace

hipError_t launch_extract_plane(const DeviceFields &f, int snap, int comp, double *out,
                                hipStream_t stream)
{
    if (comp < 0 || comp > 5 || snap < 0 || (f.nsnap == 0 && snap > 0) ||
        (f.nsnap > 0 && snap >= f.nsnap))
        return hipErrorInvalidValue;
    const size_t esz = f.prec == PTC_FP32 ? sizeof(float) : sizeof(double);
    const char *data = (const char *)f.data + (size_t)snap * (size_t)f.sstride * esz;
    int64_t blocks = (f.nodes + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(extract_plane_kernel, dim3((unsigned)blocks), dim3(256), 0, stream,
                       (const void *)data, f.layout, f.prec, f.dm[0], f.dm[1], f.dm[2], f.nodes,
                       comp, out);
    return hipGetLastError();
}

static hipError_t launch_integrate_lanes(int method, const DeviceFields &f, const StepParams &sp,
                                         const DeviceParticles &p, double t_stop,
                                         const LaunchConfig &lc, hipStream_t stream);
static hipError_t launch_integrate_park(int method, const DeviceFields &f, const StepParams &sp,
                                        const DeviceParticles &p, double t_stop,
                                        const LaunchConfig &lc, hipStream_t stream);

hipError_t launch_integrate(int method, const DeviceFields &f, const StepParams &sp,
                            const DeviceParticles &p, double t_stop, const LaunchConfig &lc,
                            hipStream_t stream)
{
    if (f.prec == PTC_FP32) return launch_integrate_mixed(method, f, sp, p, t_stop, lc, stream);
    if (f.layout == LAYOUT_PAIRS || f.layout == LAYOUT_COLUMNS) {
        // FP64 pair or column rows: PI controller, the trio or group kernel
        // only (a snapshot series in column rows as well)
        if (lc.ctrl != CTRL_PI || lc.tail != kTailGroupOnly) return hipErrorInvalidValue;
        const int g = sp.pow2 ? 1 : 0;
        if (f.nsnap > 0) {
            if (f.layout != LAYOUT_COLUMNS) return hipErrorInvalidValue;
#define GROUP_TC(M, MO, G)                                                     \
    if (method == M && f.mode == MO && g == G)                                 \
        return lc.lanes == 3                                                   \
                   ? launch_trio_t<M, MO, G, PTC_FP64, ROWS_COLUMNS, CTRL_PI, 1>(f, sp, p, t_stop, kBlock, stream) \
                   : launch_group_t<M, MO, G, PTC_FP64, 1, CTRL_PI, ROWS_COLUMNS>(f, sp, p, t_stop, kBlock, stream);
            GROUP_TC(PTC_RK5, PTC_MODE_VB, 0) GROUP_TC(PTC_RK5, PTC_MODE_EB, 0)
            GROUP_TC(PTC_DOP853, PTC_MODE_VB, 0) GROUP_TC(PTC_DOP853, PTC_MODE_EB, 0)
            GROUP_TC(PTC_RK5, PTC_MODE_VB, 1) GROUP_TC(PTC_RK5, PTC_MODE_EB, 1)
            GROUP_TC(PTC_DOP853, PTC_MODE_VB, 1) GROUP_TC(PTC_DOP853, PTC_MODE_EB, 1)
#undef GROUP_TC
            return hipErrorInvalidValue;
        }
        const int rows = f.layout == LAYOUT_COLUMNS ? ROWS_COLUMNS : ROWS_PAIRS;
#define GROUP_R(M, MO, G, R)                                                   \
    if (method == M && f.mode == MO && g == G && rows == R)                    \
        return lc.lanes == 3                                                   \
                   ? launch_trio_t<M, MO, G, PTC_FP64, R>(f, sp, p, t_stop, kBlock, stream) \
                   : launch_group_t<M, MO, G, PTC_FP64, 0, CTRL_PI, R>(f, sp, p, t_stop, kBlock, stream);
#define GROUP_P(M, MO, G) GROUP_R(M, MO, G, ROWS_PAIRS) GROUP_R(M, MO, G, ROWS_COLUMNS)
        GROUP_P(PTC_RK5, PTC_MODE_VB, 0) GROUP_P(PTC_RK5, PTC_MODE_EB, 0)
        GROUP_P(PTC_DOP853, PTC_MODE_VB, 0) GROUP_P(PTC_DOP853, PTC_MODE_EB, 0)
        GROUP_P(PTC_RK5, PTC_MODE_VB, 1) GROUP_P(PTC_RK5, PTC_MODE_EB, 1)
        GROUP_P(PTC_DOP853, PTC_MODE_VB, 1) GROUP_P(PTC_DOP853, PTC_MODE_EB, 1)
#undef GROUP_P
#undef GROUP_R
        return hipErrorInvalidValue;
    }
    if (f.nsnap > 0) {
        // snapshot series: node rows, PI controller; the cell cache keeps
        // the second snapshot in LDS where the device has room for it
        if (f.layout != LAYOUT_AOS || lc.ctrl != CTRL_PI) return hipErrorInvalidValue;
        if (lc.tail == kTailGroupOnly) {
            const int g = sp.pow2 ? 1 : 0;
#define GROUP_T(M, MO, G)                                                      \
    if (method == M && f.mode == MO && g == G)                                 \
        return lc.lanes == 3                                                   \
                   ? launch_trio_t<M, MO, G, PTC_FP64, ROWS_NODES, CTRL_PI, 1>(f, sp, p, t_stop, kBlock, stream) \
                   : launch_group_t<M, MO, G, PTC_FP64, 1>(f, sp, p, t_stop, kBlock, stream);
            GROUP_T(PTC_RK5, PTC_MODE_VB, 0) GROUP_T(PTC_RK5, PTC_MODE_EB, 0)
            GROUP_T(PTC_DOP853, PTC_MODE_VB, 0) GROUP_T(PTC_DOP853, PTC_MODE_EB, 0)
            GROUP_T(PTC_RK5, PTC_MODE_VB, 1) GROUP_T(PTC_RK5, PTC_MODE_EB, 1)
            GROUP_T(PTC_DOP853, PTC_MODE_VB, 1) GROUP_T(PTC_DOP853, PTC_MODE_EB, 1)
#undef GROUP_T
            return hipErrorInvalidValue;
        }
        const int c = lc.cell_cache == 1 && series_cache_fits(PTC_FP64, lc.block) ? 1 : 0;
#define SERIES(M, MO, C)                                                       \
    if (method == M && f.mode == MO && c == C)                                 \
        return launch_t<M, MO, LAYOUT_AOS, C, PTC_FP64, CTRL_PI, 0, 1>(f, sp, p, t_stop, lc, stream);
        SERIES(PTC_RK5, PTC_MODE_VB, 0) SERIES(PTC_RK5, PTC_MODE_EB, 0)
        SERIES(PTC_DOP853, PTC_MODE_VB, 0) SERIES(PTC_DOP853, PTC_MODE_EB, 0)
        SERIES(PTC_RK5, PTC_MODE_VB, 1) SERIES(PTC_RK5, PTC_MODE_EB, 1)
        SERIES(PTC_DOP853, PTC_MODE_VB, 1) SERIES(PTC_DOP853, PTC_MODE_EB, 1)
#undef SERIES
        return hipErrorInvalidValue;
    }
    // FP64 static field in node rows with the PI controller: the trio kernel
    // (three lanes per particle) or the group kernel (six) takes the tail of
    // the launch, or all of it
    if (lc.tail != 0 && f.layout == LAYOUT_AOS && lc.ctrl == CTRL_PI) {
        const int g = sp.pow2 ? 1 : 0;
        DeviceParticles q = p;
        q.park = lc.tail > 0 ? lc.tail : 0;
        if (lc.tail != kTailGroupOnly) {
            const hipError_t e = launch_integrate_park(method, f, sp, q, t_stop, lc, stream);
            if (e != hipSuccess) return e;
        }
#define GROUP(M, MO, G)                                                        \
    if (method == M && f.mode == MO && g == G)                                 \
        return lc.lanes == 3                                                   \
                   ? launch_trio_t<M, MO, G, PTC_FP64, ROWS_NODES>(f, sp, q, t_stop, kBlock, stream) \
                   : launch_group_t<M, MO, G, PTC_FP64>(f, sp, q, t_stop, kBlock, stream);
        GROUP(PTC_RK5, PTC_MODE_VB, 0) GROUP(PTC_RK5, PTC_MODE_EB, 0)
        GROUP(PTC_DOP853, PTC_MODE_VB, 0) GROUP(PTC_DOP853, PTC_MODE_EB, 0)
        GROUP(PTC_RK5, PTC_MODE_VB, 1) GROUP(PTC_RK5, PTC_MODE_EB, 1)
        GROUP(PTC_DOP853, PTC_MODE_VB, 1) GROUP(PTC_DOP853, PTC_MODE_EB, 1)
#undef GROUP
        return hipErrorInvalidValue;
    }
    DeviceParticles q = p;
    q.park = 0;
    return launch_integrate_lanes(method, f, sp, q, t_stop, lc, stream);
}

// the lane-per-particle kernels that hand their last particles over
static hipError_t launch_integrate_park(int method, const DeviceFields &f, const StepParams &sp,
                                        const DeviceParticles &p, double t_stop,
                                        const LaunchConfig &lc, hipStream_t stream)
{
    const int geom = sp.pow2 && method == PTC_DOP853 && lc.cell_cache == 1 ? 1 : 0;
#define PARKED(M, MO, C, G)                                                    \
    if (method == M && f.mode == MO && lc.cell_cache == C && geom == G)        \
        return launch_t<M, MO, LAYOUT_AOS, C, PTC_FP64, CTRL_PI, G, 0, 1>(f, sp, p, t_stop, lc, \
                                                                          stream);
    PARKED(PTC_RK5, PTC_MODE_VB, 1, 0) PARKED(PTC_RK5, PTC_MODE_EB, 1, 0)
    PARKED(PTC_DOP853, PTC_MODE_VB, 1, 0) PARKED(PTC_DOP853, PTC_MODE_EB, 1, 0)
    PARKED(PTC_DOP853, PTC_MODE_VB, 1, 1) PARKED(PTC_DOP853, PTC_MODE_EB, 1, 1)
    PARKED(PTC_RK5, PTC_MODE_VB, 0, 0) PARKED(PTC_RK5, PTC_MODE_EB, 0, 0)
    PARKED(PTC_DOP853, PTC_MODE_VB, 0, 0) PARKED(PTC_DOP853, PTC_MODE_EB, 0, 0)
#undef PARKED
    return hipErrorInvalidValue;
}

// the lane-per-particle kernels of a static FP64 field
static hipError_t launch_integrate_lanes(int method, const DeviceFields &f, const StepParams &sp,
                                         const DeviceParticles &p, double t_stop,
                                         const LaunchConfig &lc, hipStream_t stream)
{
    // FP64 power-of-two geometry (box sizes and velocity unit powers of two:
    // the six divisions by them become exact multiplications): DOP853 with
    // the cell cache and the PI controller only, where it spills less and
    // runs 1.07x faster; RK5 spills four times as much with it.
    const int geom = sp.pow2 && method == PTC_DOP853 && f.layout == LAYOUT_AOS &&
                             lc.cell_cache == 1 && lc.ctrl == CTRL_PI
                         ? 1
                         : 0;
#define DISPATCH(M, MO, L, C, CT, G)                                           \
    if (method == M && f.mode == MO && f.layout == L && lc.cell_cache == C &&  \
        lc.ctrl == CT && geom == G)                                            \
        return launch_t<M, MO, L, C, PTC_FP64, CT, G>(f, sp, p, t_stop, lc, stream);
#define DISPATCH_MODES(L, C, CT, G)                                            \
    DISPATCH(PTC_RK5, PTC_MODE_VB, L, C, CT, G)                                \
    DISPATCH(PTC_RK5, PTC_MODE_EB, L, C, CT, G)                                \
    DISPATCH(PTC_DOP853, PTC_MODE_VB, L, C, CT, G)                             \
    DISPATCH(PTC_DOP853, PTC_MODE_EB, L, C, CT, G)
    DISPATCH_MODES(LAYOUT_AOS, 1, CTRL_PI, 0)
    DISPATCH(PTC_DOP853, PTC_MODE_VB, LAYOUT_AOS, 1, CTRL_PI, 1)
    DISPATCH(PTC_DOP853, PTC_MODE_EB, LAYOUT_AOS, 1, CTRL_PI, 1)
    DISPATCH_MODES(LAYOUT_AOS, 1, CTRL_RECORD, 0)
    DISPATCH_MODES(LAYOUT_AOS, 1, CTRL_LEARNED, 0)
    DISPATCH_MODES(LAYOUT_AOS, 0, CTRL_PI, 0)
    DISPATCH_MODES(LAYOUT_SOA, 0, CTRL_PI, 0)
    DISPATCH_MODES(LAYOUT_BRICK, 0, CTRL_PI, 0)
#undef DISPATCH_MODES
#undef DISPATCH
    return hipErrorInvalidValue;
}

hipError_t launch_probe(const DeviceFields &f, const StepParams &sp, int op, int64_t n,
                        const double *in, double *out, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    if (op < 0 || op > 7) return hipErrorInvalidValue;
    if (op >= 5 && f.layout != LAYOUT_AOS) return hipErrorInvalidValue;
    if (rows_layout(f.layout)) return hipErrorInvalidValue;  // the probes read node rows
    if (f.prec == PTC_FP32) return launch_probe_mixed(f, sp, op, n, in, out, stream);
    if (f.mode == PTC_MODE_VB && f.layout == LAYOUT_SOA)
        return launch_probe_t<PTC_MODE_VB, LAYOUT_SOA>(f, sp, op, n, in, out, stream);
    if (f.mode == PTC_MODE_VB && f.layout == LAYOUT_AOS)
        return launch_probe_t<PTC_MODE_VB, LAYOUT_AOS>(f, sp, op, n, in, out, stream);
    if (f.mode == PTC_MODE_EB && f.layout == LAYOUT_SOA)
        return launch_probe_t<PTC_MODE_EB, LAYOUT_SOA>(f, sp, op, n, in, out, stream);
    if (f.mode == PTC_MODE_EB && f.layout == LAYOUT_AOS)
        return launch_probe_t<PTC_MODE_EB, LAYOUT_AOS>(f, sp, op, n, in, out, stream);
    if (f.mode == PTC_MODE_VB)
        return launch_probe_t<PTC_MODE_VB, LAYOUT_BRICK>(f, sp, op, n, in, out, stream);
    return launch_probe_t<PTC_MODE_EB, LAYOUT_BRICK>(f, sp, op, n, in, out, stream);
}

hipError_t launch_to_bricks(const DeviceFields &f, hipStream_t stream)
{
    if (f.prec != PTC_FP64 || !f.bricks) return hipErrorInvalidValue;
    hipLaunchKernelGGL(to_bricks_kernel, dim3(65536), dim3(256), 0, stream, (const double *)f.data,
                       f.dm[0], f.dm[1], f.dm[2], f.nodes, (double *)f.bricks);
    return hipGetLastError();
}

hipError_t launch_totals(const DeviceParticles &p, int64_t *d_out, hipStream_t stream)
{
    hipError_t e = launch_zero_i64(d_out, 5, stream);
    if (e != hipSuccess) return e;
    int64_t blocks = (p.np + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(totals_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, p,
                       (unsigned long long *)d_out);
    return hipGetLastError();
}

hipError_t launch_zero_i32(int32_t *ptr, int64_t n, hipStream_t stream)
{
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(zero_i32_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, ptr, n);
    return hipGetLastError();
}

hipError_t launch_clear_flag(int32_t *flags, int64_t n, int32_t bit, hipStream_t stream)
{
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(clear_flag_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, flags, n, bit);
    return hipGetLastError();
}

hipError_t launch_zero_i64(int64_t *ptr, int64_t n, hipStream_t stream)
{
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(zero_i64_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, ptr, n);
    return hipGetLastError();
}

hipError_t launch_to_pairs_f32(const double *soa, const int dm[3], float *out, hipStream_t stream)
{
    const int64_t nodes = (int64_t)dm[0] * dm[1] * dm[2];
    int64_t blocks = (nodes + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(to_pairs_f32_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, soa,
                       dm[0], nodes, out);
    return hipGetLastError();
}

hipError_t launch_convert_rows(const DeviceFields &f, void *out, int layout, int prec,
                               hipStream_t stream)
{
    if (f.nsnap > 0 || f.layout == LAYOUT_BRICK || layout == LAYOUT_SOA ||
        layout == LAYOUT_BRICK || (layout == LAYOUT_QUADS && prec != PTC_FP32))
        return hipErrorInvalidValue;
    int64_t blocks = (f.nodes + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(convert_rows_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, f, out,
                       layout, prec);
    return hipGetLastError();
}

hipError_t launch_to_pairs_f64(const double *soa, const int dm[3], double *out, hipStream_t stream)
{
    const int64_t nodes = (int64_t)dm[0] * dm[1] * dm[2];
    int64_t blocks = (nodes + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(to_pairs_f64_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, soa,
                       dm[0], nodes, out);
    return hipGetLastError();
}

hipError_t launch_to_quads_f32(const double *soa, const int dm[3], float *out, hipStream_t stream)
{
    const int64_t nodes = (int64_t)dm[0] * dm[1] * dm[2];
    int64_t blocks = (nodes + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(to_quads_f32_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, soa,
                       dm[0], dm[1], nodes, out);
    return hipGetLastError();
}

hipError_t launch_to_columns_f32(const double *soa, const int dm[3], float *out, hipStream_t stream)
{
    const int64_t nodes = (int64_t)dm[0] * dm[1] * dm[2];
    int64_t blocks = (nodes + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(to_columns_f32_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, soa,
                       dm[0], dm[1], dm[2], nodes, out);
    return hipGetLastError();
}

hipError_t launch_to_columns_f64(const double *soa, const int dm[3], double *out, hipStream_t stream)
{
    const int64_t nodes = (int64_t)dm[0] * dm[1] * dm[2];
    int64_t blocks = (nodes + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(to_columns_f64_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, soa,
                       dm[0], dm[1], dm[2], nodes, out);
    return hipGetLastError();
}

hipError_t launch_to_aos(const double *soa, int64_t nodes, void *aos, int prec, hipStream_t stream)
{
    int64_t blocks = (nodes + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (prec == PTC_FP32) {
        hipLaunchKernelGGL(to_nodes_f32_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, soa,
                           nodes, (float *)aos);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(to_aos_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, soa, nodes,
                       (double *)aos);
    return hipGetLastError();
}

}  // namespace ptc
