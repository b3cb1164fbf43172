// Device code of the integrator, shared by the FP64 translation unit
// (integrator.hip, no FMA contraction: bitwise parity with the reference)
// and the mixed-precision one (integrator_mixed.hip).  See integrator.hip
// for the design.
#pragma once
#include "internal.hpp"
#include "libm_pow.hpp"
#include "mlp_device.hpp"
#include "particles_amd.h"
#include <type_traits>

namespace ptc {
namespace {

constexpr int kBlock = 256;

// ---------------------------------------------------------------------------
// trilinear gather, sources/interpolations.F90:104-226
// ---------------------------------------------------------------------------

// i == n wraps to 0 (see corners()); anything else outside [0, n) can only
// come from a non-finite position and is clamped to 0 so the gather stays in
// bounds (the non-finite result then retires the particle).
__device__ __forceinline__ int wrap_idx(int i, int n)
{
    i = i >= n ? i - n : i;
    return (unsigned)i < (unsigned)n ? i : 0;
}

template <class W>
struct CornersT {
    uint32_t o[8];  // node offsets of the 8 corners, reference order
    W w[8];         // trilinear weights, reference order
    bool adjacent;  // right index == left + 1 (periodic) in every direction
};
using Corners = CornersT<double>;

// Index and weight math of interpolations.F90:140-149 and :207-210.  The
// right index is its own floor of x + 1/dm, as in the reference, so at a
// cell face it can differ from left+1; an index equal to dm (x just below a
// periodic image, where the reference reads one past the array) wraps to 0.
// W = float (mixed precision): the in-cell fractions are formed in FP64 as
// above and rounded once; the weight products are FP32.
// GEOM = 1: every xs is a power of two, so (pos - xl) / xs equals
// (pos - xl) * (1 / xs) bit for bit and the division is skipped.
template <class W, int GEOM = 0>
__device__ __forceinline__ void corners(const StepParams &P, const double *pos, CornersT<W> &c,
                                        float *frac = nullptr)
{
    int l[3], r[3];
    double xd[3], xu[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double x = GEOM ? (pos[i] - P.xl[i]) * P.rxs[i] : (pos[i] - P.xl[i]) / P.xs[i];
        const double xq = x + P.inv_dm[i];
        double xp = P.dmd[i] * (xq - floor(xq));
        r[i] = wrap_idx((int)floor(xp), P.dm[i]);
        xp = P.dmd[i] * (x - floor(x));
        const double pl = floor(xp);
        xd[i] = xp - pl;
        xu[i] = 1.0 - xd[i];
        l[i] = wrap_idx((int)pl, P.dm[i]);
        if (frac) frac[i] = float(xd[i]);
    }
    c.adjacent = true;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const int lp = l[i] + 1 == P.dm[i] ? 0 : l[i] + 1;
        c.adjacent = c.adjacent && r[i] == lp;
    }
    const uint32_t sy = (uint32_t)P.dm[0], sz = (uint32_t)P.dm[0] * (uint32_t)P.dm[1];
    const uint32_t yl = (uint32_t)l[1] * sy, yr = (uint32_t)r[1] * sy;
    const uint32_t zl = (uint32_t)l[2] * sz, zr = (uint32_t)r[2] * sz;
    const uint32_t xl = (uint32_t)l[0], xr = (uint32_t)r[0];
    c.o[0] = xl + yl + zl; c.o[1] = xl + yl + zr; c.o[2] = xl + yr + zl; c.o[3] = xl + yr + zr;
    c.o[4] = xr + yl + zl; c.o[5] = xr + yl + zr; c.o[6] = xr + yr + zl; c.o[7] = xr + yr + zr;
    const W u0 = W(xu[0]), u1 = W(xu[1]), u2 = W(xu[2]);
    const W d0 = W(xd[0]), d1 = W(xd[1]), d2 = W(xd[2]);
    c.w[0] = u0 * u1 * u2;
    c.w[1] = u0 * u1 * d2;
    c.w[2] = u0 * d1 * u2;
    c.w[3] = u0 * d1 * d2;
    c.w[4] = d0 * u1 * u2;
    c.w[5] = d0 * u1 * d2;
    c.w[6] = d0 * d1 * u2;
    c.w[7] = d0 * d1 * d2;
}

// f = w0 v0 + w1 v1 + ... + w7 v7, summed left to right per component
template <class T>
__device__ __forceinline__ void accumulate_row(int j, T w, const T *v, T *f)
{
    if (j == 0) {
#pragma unroll
        for (int k = 0; k < 6; k++) f[k] = w * v[k];
    } else {
#pragma unroll
        for (int k = 0; k < 6; k++) f[k] = f[k] + w * v[k];
    }
}

// eight 48-byte node rows of the AOS array
__device__ __forceinline__ void gather_rows(const double *__restrict__ d, const Corners &c, double *f)
{
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const double2 *__restrict__ q = (const double2 *)(d + (size_t)c.o[j] * 6);
        const double2 a = q[0], b = q[1], e = q[2];
        const double v[6] = {a.x, a.y, b.x, b.y, e.x, e.y};
        accumulate_row(j, c.w[j], v, f);
    }
}

// f[0..2] = U (or E), f[3..5] = B; each a left-to-right weighted sum of the
// eight corners (interpolations.F90:212-220).
template <int LAYOUT>
__device__ __forceinline__ void gather(const DeviceFields &F, const StepParams &P, const double *pos,
                                       double *f)
{
    Corners c;
    corners(P, pos, c);
    const double *__restrict__ d = (const double *)F.data;
    if constexpr (LAYOUT == LAYOUT_SOA) {
        const int64_t N = F.nodes;
#pragma unroll
        for (int k = 0; k < 6; k++) {
            const double *__restrict__ p = d + k * N;
            double s = c.w[0] * p[c.o[0]];
#pragma unroll
            for (int j = 1; j < 8; j++) s = s + c.w[j] * p[c.o[j]];
            f[k] = s;
        }
    } else if constexpr (LAYOUT == LAYOUT_AOS) {
        gather_rows(d, c, f);
    } else {
        if (c.adjacent) {
            // the cell's own 384-byte block: row j holds corner j
            const double2 *__restrict__ q =
                (const double2 *)((const double *)F.bricks + (size_t)c.o[0] * 48);
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const double2 a = q[3 * j], b = q[3 * j + 1], e = q[3 * j + 2];
                const double v[6] = {a.x, a.y, b.x, b.y, e.x, e.y};
                accumulate_row(j, c.w[j], v, f);
            }
        } else {
            gather_rows(d, c, f);
        }
    }
}

// ---------------------------------------------------------------------------
// relativistic Lorentz force, sources/accelerations.F90:59-96
// ---------------------------------------------------------------------------
// GEOM = 1: vun is a power of two as well (v / vun == v * (1 / vun) exactly)
template <int MODE, int GEOM = 0>
__device__ __forceinline__ void force_from_fields(const StepParams &P, const double *s,
                                                  const double *f, double *a)
{
    const double g = sqrt(s[3] * s[3] + s[4] * s[4] + s[5] * s[5] + 1.0);
    const double v0 = s[3] / g, v1 = s[4] / g, v2 = s[5] / g;
    a[0] = GEOM ? v0 * P.rvun : v0 / P.vun;
    a[1] = GEOM ? v1 * P.rvun : v1 / P.vun;
    a[2] = GEOM ? v2 * P.rvun : v2 / P.vun;
    if constexpr (MODE == PTC_MODE_EB) {
        a[3] = P.qom * (f[0] + (v1 * f[5] - v2 * f[4]));
        a[4] = P.qom * (f[1] + (v2 * f[3] - v0 * f[5]));
        a[5] = P.qom * (f[2] + (v0 * f[4] - v1 * f[3]));
    } else {
        const double w0 = v0 - f[0], w1 = v1 - f[1], w2 = v2 - f[2];
        a[3] = P.qom * (w1 * f[5] - w2 * f[4]);
        a[4] = P.qom * (w2 * f[3] - w0 * f[5]);
        a[5] = P.qom * (w0 * f[4] - w1 * f[3]);
    }
}

// Mixed precision: the same force in FP32 from the FP64 stage point.
template <int MODE>
__device__ __forceinline__ void force_from_fields_f32(const StepParams &P, const double *s,
                                                      const float *f, float *a)
{
    // hardware reciprocal square root (1 ulp) of gamma squared in place of
    // the square root and the IEEE divisions
    const float ux = float(s[3]), uy = float(s[4]), uz = float(s[5]);
    const float rg = __builtin_amdgcn_rsqf(ux * ux + uy * uy + uz * uz + 1.0f);
    const float v0 = ux * rg, v1 = uy * rg, v2 = uz * rg;
    const float rvun = float(P.rvun), qom = float(P.qom);
    a[0] = v0 * rvun;
    a[1] = v1 * rvun;
    a[2] = v2 * rvun;
    if constexpr (MODE == PTC_MODE_EB) {
        a[3] = qom * (f[0] + (v1 * f[5] - v2 * f[4]));
        a[4] = qom * (f[1] + (v2 * f[3] - v0 * f[5]));
        a[5] = qom * (f[2] + (v0 * f[4] - v1 * f[3]));
    } else {
        const float w0 = v0 - f[0], w1 = v1 - f[1], w2 = v2 - f[2];
        a[3] = qom * (w1 * f[5] - w2 * f[4]);
        a[4] = qom * (w2 * f[3] - w0 * f[5]);
        a[5] = qom * (w0 * f[4] - w1 * f[3]);
    }
}

// FP32 nodes: eight floats (U or E x,y,z, B x,y,z, |B|, 0), 32 bytes, so a
// corner is one 16-byte and one 8-byte load
__device__ __forceinline__ void load_node_f32(const float *__restrict__ d, uint32_t o, float *v)
{
    const float4 a = *(const float4 *)(d + (size_t)o * 8);
    const float2 b = *(const float2 *)(d + (size_t)o * 8 + 4);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y;
}

template <int MODE, int GEOM = 0>
__device__ __forceinline__ void rhs_f32(const DeviceFields &F, const StepParams &P, const double *s,
                                        float *a)
{
    CornersT<float> c;
    corners<float, GEOM>(P, s, c);
    const float *__restrict__ d = (const float *)F.data;
    float f[6];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        float v[6];
        load_node_f32(d, c.o[j], v);
        accumulate_row(j, c.w[j], v, f);
    }
    force_from_fields_f32<MODE>(P, s, f, a);
}

template <int MODE, int LAYOUT>
__device__ __forceinline__ void rhs(const DeviceFields &F, const StepParams &P, const double *s,
                                    double *a)
{
    double f[6];
    gather<LAYOUT>(F, P, s, f);
    force_from_fields<MODE>(P, s, f, a);
}

// The RHS as the integrator calls it.  With a cell cache, each lane keeps the
// 48 corner values of the last cell it gathered from (the node rows are
// identified by the offsets of corners 0 and 7, which fix both the left and
// the right index triples); a stage point in the same cell re-weights those
// values instead of loading them again.  The values and the summation order
// are the same either way, so the result is bit-identical.  CACHE 0 reloads
// every time (SOA and brick layouts).
// PREC = PTC_FP32 is the mixed-precision force: FP32 nodes, weights and
// arithmetic, slopes returned as float (K = float).
// TIME = 1: the fields are a series of snapshots (F.nsnap, F.stime); the
// caller sets the stage time with at() before each evaluation, and the
// fields there are f_k + a (f_{k+1} - f_k), the two bracketing snapshots'
// trilinear interpolants blended linearly in time (node rows).  With the
// cell cache, snapshot k's 48 corner values sit in registers as for a static
// field and snapshot k+1's in LDS (FP64) or registers too (FP32): in LDS
// each wave owns a 48 x 64 slab, value j of lane l at slab[64 j + l], so a
// read of one value by the wave covers 64 consecutive elements (no bank
// conflicts) and the 48 reads differ by compile-time offsets from one
// per-lane address (no block size in the address); the key adds the
// interval.  With TIME = 0, at() does nothing and the stage
// times are never formed.
// SPLIT > 0 (static FP64 field, PI controller): the last SPLIT of the 48
// cached values live in LDS (lds1, the per-wave slab layout below) instead
// of registers.
// Adds a wave's cell loads to the batch's counter (one atomic per wave, at
// the end of a launch): every lane passes its own count, zero for lanes
// that do not stand for a particle of their own.
__device__ __forceinline__ void add_cell_loads(const DeviceParticles &S, uint32_t loads)
{
    unsigned long long l = loads;
    for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off);
    if (__lane_id() == 0 && S.loads && l) atomicAdd(S.loads, l);
}

template <int MODE, int LAYOUT, int CACHE, int PREC, int GEOM = 0, int TIME = 0, int SPLIT = 0>
struct Force {
    using K = std::conditional_t<PREC == PTC_FP32, float, double>;
    static constexpr int kRegVals = CACHE == 1 ? 48 - SPLIT : 1;
    const DeviceFields &F;
    const StepParams &P;
    uint32_t key0 = 0xffffffffu, key7 = 0xffffffffu;
    uint32_t loads = 0;  // cell loads (eight corner rows of one snapshot) by this lane
    K v[kRegVals > 0 ? kRegVals : 1];
    double ts = 0.0;  // TIME: stage time of the next evaluation
    int kh = 0;       // TIME: the lane's last snapshot interval, in [0, nsnap - 2]
    int kc = -1;      // TIME with the cache: 2 k + (snapshot k+1 loaded) of the cached cell
    K *lds1 = nullptr;  // TIME with the cache: this lane's snapshot-k+1 column (stride 64)
    // TIME with the cache in FP32: snapshot k+1 in registers instead (the
    // FP32 kernel has room for 48 more floats; FP64 does not)
    static constexpr bool kSecondInRegs =
        TIME == 1 && CACHE == 1 && PREC == PTC_FP32;
    K v1[kSecondInRegs ? 48 : 1];

    __device__ __forceinline__ Force(const DeviceFields &F_, const StepParams &P_) : F(F_), P(P_) {}

    __device__ __forceinline__ void at(double t)
    {
        if constexpr (TIME == 1) ts = t;
    }

    // Snapshot interval of the stage time: k with stime[k] <= ts <
    // stime[k+1] and a = (ts - stime[k]) / (stime[k+1] - stime[k]) (two =
    // true).  At or before the first time (or a NaN time) snapshot 0 alone,
    // at or after the last time the last snapshot alone (two = false).  The
    // search starts from the lane's last interval (stage times move forward
    // by less than a snapshot spacing, so that or the next one nearly
    // always holds) and bisects otherwise (a lane that picked up a particle
    // at another time).  Every index stays in [0, nsnap - 1] because
    // stime[0] < ts < stime[nsnap - 1].
    __device__ __forceinline__ int interval(double &a, bool &two)
    {
        const int ns = F.nsnap;
        const double *__restrict__ st = F.stime;
        a = 0.0;
        two = false;
        if (ns <= 1 || !(ts > st[0])) return 0;
        if (ts >= st[ns - 1]) return ns - 1;
        int k = kh;
        if (!(st[k] <= ts && ts < st[k + 1])) {
            if (k + 2 < ns && st[k + 1] <= ts && ts < st[k + 2]) {
                k = k + 1;
            } else {
                int lo = 0, hi = ns - 1;  // st[lo] <= ts < st[hi]
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (st[mid] <= ts) lo = mid;
                    else hi = mid;
                }
                k = lo;
            }
        }
        kh = k;
        a = (ts - st[k]) / (st[k + 1] - st[k]);
        two = true;
        return k;
    }

    // TIME with the cache: loads the corner rows of snapshot k (registers)
    // and, if two, of snapshot k+1 (LDS), unless the cell and interval are
    // cached.  A pair cached for interval k also serves snapshot k alone.
    __device__ __forceinline__ void fill_pair(const uint32_t *o, int k, bool two)
    {
        const int key = 2 * k + (two ? 1 : 0);
        if (o[0] == key0 && o[7] == key7 && (key == kc || (!two && key + 1 == kc))) return;
        loads += two ? 2u : 1u;
        if constexpr (PREC == PTC_FP32) {
            const float *__restrict__ d = (const float *)F.data + (size_t)k * F.sstride;
#pragma unroll
            for (int j = 0; j < 8; j++) load_node_f32(d, o[j], &v[6 * j]);
            if (two) {
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    float w[6];
                    load_node_f32(d + F.sstride, o[j], w);
#pragma unroll
                    for (int i = 0; i < 6; i++) {
                        if constexpr (kSecondInRegs) v1[6 * j + i] = w[i];
                        else lds1[(6 * j + i) * 64] = w[i];
                    }
                }
            }
        } else {
            const double *__restrict__ d = (const double *)F.data + (size_t)k * F.sstride;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const double2 *__restrict__ q = (const double2 *)(d + (size_t)o[j] * 6);
                const double2 x0 = q[0], x1 = q[1], x2 = q[2];
                v[6 * j + 0] = x0.x; v[6 * j + 1] = x0.y; v[6 * j + 2] = x1.x;
                v[6 * j + 3] = x1.y; v[6 * j + 4] = x2.x; v[6 * j + 5] = x2.y;
            }
            if (two) {
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const double2 *__restrict__ q =
                        (const double2 *)(d + F.sstride + (size_t)o[j] * 6);
                    const double2 x0 = q[0], x1 = q[1], x2 = q[2];
                    const double w[6] = {x0.x, x0.y, x1.x, x1.y, x2.x, x2.y};
#pragma unroll
                    for (int i = 0; i < 6; i++) {
                        if constexpr (kSecondInRegs) v1[6 * j + i] = w[i];
                        else lds1[(6 * j + i) * 64] = w[i];
                    }
                }
            }
        }
        key0 = o[0];
        key7 = o[7];
        kc = key;
    }

    // TIME: the interpolated fields (U or E, B) at s and the stage time
    __device__ __forceinline__ void fields_at(const double *s, K *f)
    {
        static_assert(TIME == 1 && LAYOUT == LAYOUT_AOS, "a snapshot series is read as node rows");
        CornersT<K> c;
        corners<K, GEOM>(P, s, c);
        double a;
        bool two;
        const int k = interval(a, two);
        if constexpr (CACHE == 1) {
            fill_pair(c.o, k, two);
#pragma unroll
            for (int j = 0; j < 8; j++) accumulate_row(j, c.w[j], &v[6 * j], f);
            if (two) {
                K g[6];
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    K w[6];
#pragma unroll
                    for (int i = 0; i < 6; i++) {
                        if constexpr (kSecondInRegs) w[i] = v1[6 * j + i];
                        else w[i] = lds1[(6 * j + i) * 64];
                    }
                    accumulate_row(j, c.w[j], w, g);
                }
                const K ak = K(a);
#pragma unroll
                for (int i = 0; i < 6; i++) f[i] = f[i] + ak * (g[i] - f[i]);
            }
        } else if constexpr (PREC == PTC_FP32) {
            loads += two ? 2u : 1u;
            const float *__restrict__ d = (const float *)F.data + (size_t)k * F.sstride;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                float w[6];
                load_node_f32(d, c.o[j], w);
                accumulate_row(j, c.w[j], w, f);
            }
            if (two) {
                const float *__restrict__ e = d + F.sstride;
                float g[6];
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    float w[6];
                    load_node_f32(e, c.o[j], w);
                    accumulate_row(j, c.w[j], w, g);
                }
                const float af = float(a);
#pragma unroll
                for (int i = 0; i < 6; i++) f[i] = f[i] + af * (g[i] - f[i]);
            }
        } else {
            loads += two ? 2u : 1u;
            const double *__restrict__ d = (const double *)F.data + (size_t)k * F.sstride;
            gather_rows(d, c, f);
            if (two) {
                double g[6];
                gather_rows(d + F.sstride, c, g);
#pragma unroll
                for (int i = 0; i < 6; i++) f[i] = f[i] + a * (g[i] - f[i]);
            }
        }
    }

    // loads the corner rows of the cell with corner offsets o unless cached
    __device__ __forceinline__ void fill(const uint32_t *o)
    {
        static_assert(CACHE == 1 && LAYOUT == LAYOUT_AOS, "the cell cache reads AOS node rows");
        if (o[0] == key0 && o[7] == key7) return;
        loads++;
        if constexpr (PREC == PTC_FP32) {
            const float *__restrict__ d = (const float *)F.data;
#pragma unroll
            for (int j = 0; j < 8; j++) load_node_f32(d, o[j], &v[6 * j]);
        } else {
            const double *__restrict__ d = (const double *)F.data;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const double2 *__restrict__ q = (const double2 *)(d + (size_t)o[j] * 6);
                const double2 x0 = q[0], x1 = q[1], x2 = q[2];
                const double w[6] = {x0.x, x0.y, x1.x, x1.y, x2.x, x2.y};
#pragma unroll
                for (int i = 0; i < 6; i++) set_cached(6 * j + i, w[i]);
            }
        }
        key0 = o[0];
        key7 = o[7];
    }

    // cached value q (0..47): registers, or the LDS slab for the last SPLIT
    __device__ __forceinline__ K cached(int q) const
    {
        if constexpr (SPLIT > 0) {
            if (q >= kRegVals) return lds1[(q - kRegVals) * 64];
        }
        return v[q];
    }
    __device__ __forceinline__ void set_cached(int q, K x)
    {
        if constexpr (SPLIT > 0) {
            if (q >= kRegVals) {
                lds1[(q - kRegVals) * 64] = x;
                return;
            }
        }
        v[q] = x;
    }

    // the RHS at s (bit-identical to operator()), also returning the
    // interpolated fields (U or E, B) and the position inside the cell
    __device__ __forceinline__ void eval_fields(const double *s, K *a, float *ub, float *frac)
    {
        CornersT<K> c;
        corners<K, GEOM>(P, s, c, frac);
        fill(c.o);
        K f[6];
#pragma unroll
        for (int j = 0; j < 8; j++) accumulate_row(j, c.w[j], &v[6 * j], f);
#pragma unroll
        for (int k = 0; k < 6; k++) ub[k] = float(f[k]);
        if constexpr (PREC == PTC_FP32) force_from_fields_f32<MODE>(P, s, f, a);
        else force_from_fields<MODE, GEOM>(P, s, f, a);
    }

    // makes the cell around s the cached one; w = its trilinear weights,
    // frac = the position inside the cell (0..1 per axis)
    __device__ __forceinline__ void load_cell(const double *s, float *w, float *frac)
    {
        CornersT<K> c;
        corners<K, GEOM>(P, s, c, frac);
        fill(c.o);
#pragma unroll
        for (int j = 0; j < 8; j++) w[j] = float(c.w[j]);
    }

    __device__ __forceinline__ void operator()(const double *s, K *a)
    {
        if constexpr (TIME == 1) {
            K f[6];
            fields_at(s, f);
            if constexpr (PREC == PTC_FP32) force_from_fields_f32<MODE>(P, s, f, a);
            else force_from_fields<MODE, GEOM>(P, s, f, a);
        } else if constexpr (PREC == PTC_FP32) {
            static_assert(LAYOUT == LAYOUT_AOS, "FP32 fields are stored as AOS nodes");
            if constexpr (CACHE == 0) {
                loads++;
                rhs_f32<MODE, GEOM>(F, P, s, a);
            } else {
                CornersT<float> c;
                corners<float, GEOM>(P, s, c);
                fill(c.o);
                float f[6];
#pragma unroll
                for (int j = 0; j < 8; j++) accumulate_row(j, c.w[j], &v[6 * j], f);
                force_from_fields_f32<MODE>(P, s, f, a);
            }
        } else if constexpr (CACHE == 0) {
            loads++;
            rhs<MODE, LAYOUT>(F, P, s, a);
        } else {
            static_assert(LAYOUT == LAYOUT_AOS, "the cell cache reads AOS node rows");
            Corners c;
            corners<double, GEOM>(P, s, c);
            fill(c.o);
            double f[6];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                double w[6];
#pragma unroll
                for (int i = 0; i < 6; i++) w[i] = cached(6 * j + i);
                accumulate_row(j, c.w[j], w, f);
            }
            force_from_fields<MODE, GEOM>(P, s, f, a);
        }
    }
};

// ---------------------------------------------------------------------------
// step-size features (no counterpart in the reference): the local field and
// orbit seen from an accepted state, for the step-size network.  From the
// cached cell around s:
//   0 log|B|              1 log(|E| / |B|)        2..7 log(|B|_face / |B|),
//   faces x-, x+, y-, y+, z-, z+ (mean |B| over each face's four corners)
//   8 log gamma           9 cos(pitch angle)      10 log(h Omega), Omega = |qom B| / gamma
//   11 log(error ratio of the accepted step)      12 log(t_face / h), t_face =
//   time to the next cell face along the velocity  13 log(h_PI / h), the PI proposal
//   14 1 if the step followed a rejection         15 log(r_L / dx), gyro-radius / cell
// h is the accepted step.  E is the E field (EB mode) or -U x B (VB mode).
// Every feature is clamped to [-30, 30]; a non-finite one becomes 0.
// ---------------------------------------------------------------------------
constexpr int kFeatures = 16;

// |B| at corner j from its three components
__device__ __forceinline__ float corner_bmag(float b0, float b1, float b2)
{
    return __builtin_amdgcn_sqrtf(b0 * b0 + b1 * b1 + b2 * b2);
}

// From the interpolated fields ub = (U or E, B) at s (only the momentum
// s[3..5] is read), |B| at the eight corners of its cell (bc, corner order
// of the gather) and the position xd inside the cell.
template <int MODE>
__device__ __forceinline__ void features_bc(const StepParams &P, const double *s, const float *ub,
                                            const float *bc, const float *xd, double h, float err,
                                            float pi_ratio, bool rej, float *f)
{
    const float U[3] = {ub[0], ub[1], ub[2]}, B[3] = {ub[3], ub[4], ub[5]};
    float E[3];
    if constexpr (MODE == PTC_MODE_EB) {
        E[0] = U[0]; E[1] = U[1]; E[2] = U[2];
    } else {
        E[0] = -(U[1] * B[2] - U[2] * B[1]);
        E[1] = -(U[2] * B[0] - U[0] * B[2]);
        E[2] = -(U[0] * B[1] - U[1] * B[0]);
    }
    // hardware square roots, reciprocals and logarithms (about 1 ulp): the
    // features are network inputs, not part of the integration
    const float bm = __builtin_amdgcn_sqrtf(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
    const float em = __builtin_amdgcn_sqrtf(E[0] * E[0] + E[1] * E[1] + E[2] * E[2]);
    const float u[3] = {float(s[3]), float(s[4]), float(s[5])};
    const float u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const float um = __builtin_amdgcn_sqrtf(u2), g = __builtin_amdgcn_sqrtf(u2 + 1.0f);
    const float cp = (u[0] * B[0] + u[1] * B[1] + u[2] * B[2]) * __builtin_amdgcn_rcpf(um * bm);
    const float qb = fabsf(float(P.qom)) * bm;
    // time to the next cell face along each axis: cells to go / cells per time
    const float rg = __builtin_amdgcn_rcpf(g);
    float tf = 3.0e38f;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const float vel = u[i] * rg * float(P.dmd[i] / (P.xs[i] * P.vun));
        const float dist = vel > 0.0f ? 1.0f - xd[i] : xd[i];
        const float ti = dist * __builtin_amdgcn_rcpf(fabsf(vel));
        tf = ti < tf ? ti : tf;
    }
    const float lb = __logf(bm);
    f[0] = lb;
    f[1] = __logf(em) - lb;
    f[2] = __logf(0.25f * (bc[0] + bc[1] + bc[2] + bc[3])) - lb;
    f[3] = __logf(0.25f * (bc[4] + bc[5] + bc[6] + bc[7])) - lb;
    f[4] = __logf(0.25f * (bc[0] + bc[1] + bc[4] + bc[5])) - lb;
    f[5] = __logf(0.25f * (bc[2] + bc[3] + bc[6] + bc[7])) - lb;
    f[6] = __logf(0.25f * (bc[0] + bc[2] + bc[4] + bc[6])) - lb;
    f[7] = __logf(0.25f * (bc[1] + bc[3] + bc[5] + bc[7])) - lb;
    const float lh = __logf(float(h));
    f[8] = __logf(g);
    f[9] = cp;
    f[10] = lh + __logf(qb) - f[8];
    f[11] = __logf(err);
    f[12] = __logf(tf) - lh;
    f[13] = pi_ratio;
    f[14] = rej ? 1.0f : 0.0f;
    f[15] = __logf(um * __builtin_amdgcn_sqrtf(fmaxf(1.0f - cp * cp, 0.0f)) *
                   float(P.dmd[0] / P.xs[0])) - __logf(qb);
#pragma unroll
    for (int i = 0; i < kFeatures; i++) {
        float x = f[i];
        x = x > 30.0f ? 30.0f : x;
        x = x < -30.0f ? -30.0f : x;
        f[i] = x == x ? x : 0.0f;
    }
}

// The same from the cached corner rows cv (48 values) of the cell.
template <int MODE, class V>
__device__ __forceinline__ void features_core(const StepParams &P, const double *s, const float *ub,
                                              const V *cv, const float *xd, double h, float err,
                                              float pi_ratio, bool rej, float *f)
{
    float bc[8];
#pragma unroll
    for (int j = 0; j < 8; j++)
        bc[j] = corner_bmag(float(cv[6 * j + 3]), float(cv[6 * j + 4]), float(cv[6 * j + 5]));
    features_bc<MODE>(P, s, ub, bc, xd, h, err, pi_ratio, rej, f);
}

// The features at s from the cell cache (loading the cell if needed).
template <int MODE, class FORCE>
__device__ __forceinline__ void step_features(FORCE &force, const StepParams &P, const double *s,
                                              double h, float err, float pi_ratio, bool rej,
                                              float *f)
{
    float w[8], xd[3], ub[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    force.load_cell(s, w, xd);
#pragma unroll
    for (int j = 0; j < 8; j++)
#pragma unroll
        for (int k = 0; k < 6; k++) ub[k] += w[j] * float(force.v[6 * j + k]);
    features_core<MODE>(P, s, ub, force.v, xd, h, err, pi_ratio, rej, f);
}

// ---------------------------------------------------------------------------
// initial step, sources/integrations.F90:524-589 (ki = f(si) is given)
// ---------------------------------------------------------------------------
template <class K>
__device__ __forceinline__ double dti_guess(const StepParams &P, int iord, const double *si,
                                            const K *ki, const PowTables &T)
{
    double sr[6], s, q, d0, d1, d2, dti, dtf;
#pragma unroll
    for (int i = 0; i < 6; i++) sr[i] = P.atol + P.rtol * fabs(si[i]);
    s = 0.0;
#pragma unroll
    for (int i = 0; i < 6; i++) { q = si[i] / sr[i]; s = s + q * q; }
    d0 = sqrt(s / 6.0);
    s = 0.0;
#pragma unroll
    for (int i = 0; i < 6; i++) { q = double(ki[i]) / sr[i]; s = s + q * q; }
    d1 = sqrt(s / 6.0);
    dti = (fmin(d0, d1) < 1.0e-10) ? 1.0e-06 : 1.0e-01 * (d0 / d1);
    s = 0.0;
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const double sf = si[i] + dti * double(ki[i]);
        q = (sf - si[i]) / sr[i];
        s = s + q * q;
    }
    d2 = sqrt(s / 6.0) / dti;
    if (d1 > d2) d2 = d1;
    if (d2 <= 1.0e-15) {
        dtf = 1.0e-03 * dti;
        if (dtf < 1.0e-06) dtf = 1.0e-06;
    } else {
        dtf = ptc_pow_with(1.0e-02 / d2, 1.0 / (double)iord, T.log, T.exp);
    }
    dti = 1.0e+02 * dti;
    if (dtf < dti) dti = dtf;
    if (P.dtm < dti) dti = P.dtm;
    return dti;
}

// ---------------------------------------------------------------------------
// Cash-Karp RK5(4) attempt, sources/integrations.F90:131-148
// ---------------------------------------------------------------------------
// Published Cash-Karp (1990) rationals; the reference holds the same values
// over common denominators (integrations.F90:86-98).
#define CK_A21 (1.0 / 5.0)
#define CK_A31 (3.0 / 40.0)
#define CK_A32 (9.0 / 40.0)
#define CK_A41 (3.0 / 10.0)
#define CK_A42 (-9.0 / 10.0)
#define CK_A43 (6.0 / 5.0)
#define CK_A51 (-11.0 / 54.0)
#define CK_A52 (5.0 / 2.0)
#define CK_A53 (-70.0 / 27.0)
#define CK_A54 (35.0 / 27.0)
#define CK_A61 (1631.0 / 55296.0)
#define CK_A62 (175.0 / 512.0)
#define CK_A63 (575.0 / 13824.0)
#define CK_A64 (44275.0 / 110592.0)
#define CK_A65 (253.0 / 4096.0)
#define CK_C1 (37.0 / 378.0)
#define CK_C3 (250.0 / 621.0)
#define CK_C4 (125.0 / 594.0)
#define CK_C6 (512.0 / 1771.0)
#define CK_D1 (-277.0 / 64512.0)
#define CK_D3 (6925.0 / 370944.0)
#define CK_D4 (-6925.0 / 202752.0)
#define CK_D5 (-277.0 / 14336.0)
#define CK_D6 (277.0 / 7084.0)

// k1 = dt f(si) from the cached derivative f1; returns del (error / atol),
// sf, and the stage-6 slope k6 used by the dense output.  Stage slopes are
// kept in the RHS's type K (double, or float in mixed precision); stage
// points, sums and the error norm are always FP64.
// Stage nodes c_i of the Cash-Karp tableau (the row sums of its stage
// matrix): stage i is evaluated at t + c_i dt.  Used only with time-dependent
// fields; the reference's fields are static and its RHS takes no time.
#define CK_N2 (1.0 / 5.0)
#define CK_N3 (3.0 / 10.0)
#define CK_N4 (3.0 / 5.0)
#define CK_N6 (7.0 / 8.0)

template <class Rhs>
__device__ __forceinline__ double attempt_rk5(Rhs &rhsf, const StepParams &P, const double *si,
                                              const typename Rhs::K *f1, double t, double dt,
                                              double *sf, double *k6d)
{
    using K = typename Rhs::K;
    K k1[6], k2[6], k3[6], k4[6], k5[6], k6[6], a[6];
    double s[6];
#pragma unroll
    for (int c = 0; c < 6; c++) k1[c] = K(dt * double(f1[c]));
#pragma unroll
    for (int c = 0; c < 6; c++) s[c] = si[c] + CK_A21 * double(k1[c]);
    rhsf.at(t + CK_N2 * dt);
    rhsf(s, a);
#pragma unroll
    for (int c = 0; c < 6; c++) k2[c] = K(dt * double(a[c]));
#pragma unroll
    for (int c = 0; c < 6; c++) s[c] = si[c] + CK_A31 * double(k1[c]) + CK_A32 * double(k2[c]);
    rhsf.at(t + CK_N3 * dt);
    rhsf(s, a);
#pragma unroll
    for (int c = 0; c < 6; c++) k3[c] = K(dt * double(a[c]));
#pragma unroll
    for (int c = 0; c < 6; c++)
        s[c] = si[c] + CK_A41 * double(k1[c]) + CK_A42 * double(k2[c]) + CK_A43 * double(k3[c]);
    rhsf.at(t + CK_N4 * dt);
    rhsf(s, a);
#pragma unroll
    for (int c = 0; c < 6; c++) k4[c] = K(dt * double(a[c]));
#pragma unroll
    for (int c = 0; c < 6; c++)
        s[c] = si[c] + CK_A51 * double(k1[c]) + CK_A52 * double(k2[c]) + CK_A53 * double(k3[c]) +
               CK_A54 * double(k4[c]);
    rhsf.at(t + dt);
    rhsf(s, a);
#pragma unroll
    for (int c = 0; c < 6; c++) k5[c] = K(dt * double(a[c]));
#pragma unroll
    for (int c = 0; c < 6; c++)
        s[c] = si[c] + CK_A61 * double(k1[c]) + CK_A62 * double(k2[c]) + CK_A63 * double(k3[c]) +
               CK_A64 * double(k4[c]) + CK_A65 * double(k5[c]);
    rhsf.at(t + CK_N6 * dt);
    rhsf(s, a);
#pragma unroll
    for (int c = 0; c < 6; c++) k6[c] = K(dt * double(a[c]));

    const double eps = 2.2250738585072014e-308;  // tiny(1d0), integrations.F90:99
    double del = 0.0;
#pragma unroll
    for (int c = 0; c < 6; c++) {
        // sum(c(1:6) k(1:6,i)) and sum(d(1:6) k(1:6,i)), zero weights kept
        double sc = 0.0;
        sc = sc + CK_C1 * double(k1[c]);
        sc = sc + 0.0 * double(k2[c]);
        sc = sc + CK_C3 * double(k3[c]);
        sc = sc + CK_C4 * double(k4[c]);
        sc = sc + 0.0 * double(k5[c]);
        sc = sc + CK_C6 * double(k6[c]);
        sf[c] = si[c] + sc;
        double se = 0.0;
        se = se + CK_D1 * double(k1[c]);
        se = se + 0.0 * double(k2[c]);
        se = se + CK_D3 * double(k3[c]);
        se = se + CK_D4 * double(k4[c]);
        se = se + CK_D5 * double(k5[c]);
        se = se + CK_D6 * double(k6[c]);
        const double q = fabs(se / (fabs(si[c]) + fabs(double(k1[c])) + eps));
        if (c == 0 || q > del) del = q;
        k6d[c] = double(k6[c]);
    }
    return del / P.atol;
}

// ---------------------------------------------------------------------------
// DOP853 attempt, sources/integrations.F90:365-437 (Hairer-Norsett-Wanner
// coefficients, public dop853.f; same published values as the reference)
// ---------------------------------------------------------------------------
#define A0201 5.26001519587677318785587544488e-02
#define A0301 1.97250569845378994544595329183e-02
#define A0302 5.91751709536136983633785987549e-02
#define A0401 2.95875854768068491816892993775e-02
#define A0403 8.87627564304205475450678981324e-02
#define A0501 2.41365134159266685502369798665e-01
#define A0503 -8.84549479328286085344864962717e-01
#define A0504 9.24834003261792003115737966543e-01
#define A0601 3.70370370370370370370370370370e-02
#define A0604 1.70828608729473871279604482173e-01
#define A0605 1.25467687566822425016691814123e-01
#define A0701 3.71093750000000000000000000000e-02
#define A0704 1.70252211019544039314978060272e-01
#define A0705 6.02165389804559606850219397283e-02
#define A0706 -1.75781250000000000000000000000e-02
#define A0801 3.70920001185047927108779319836e-02
#define A0804 1.70383925712239993810214054705e-01
#define A0805 1.07262030446373284651809199168e-01
#define A0806 -1.53194377486244017527936158236e-02
#define A0807 8.27378916381402288758473766002e-03
#define A0901 6.24110958716075717114429577812e-01
#define A0904 -3.36089262944694129406857109825e+00
#define A0905 -8.68219346841726006818189891453e-01
#define A0906 2.75920996994467083049415600797e+01
#define A0907 2.01540675504778934086186788979e+01
#define A0908 -4.34898841810699588477366255144e+01
#define A1001 4.77662536438264365890433908527e-01
#define A1004 -2.48811461997166764192642586468e+00
#define A1005 -5.90290826836842996371446475743e-01
#define A1006 2.12300514481811942347288949897e+01
#define A1007 1.52792336328824235832596922938e+01
#define A1008 -3.32882109689848629194453265587e+01
#define A1009 -2.03312017085086261358222928593e-02
#define A1101 -9.37142430085987325717040216580e-01
#define A1104 5.18637242884406370830023853209e+00
#define A1105 1.09143734899672957818500254654e+00
#define A1106 -8.14978701074692612513997267357e+00
#define A1107 -1.85200656599969598641566180701e+01
#define A1108 2.27394870993505042818970056734e+01
#define A1109 2.49360555267965238987089396762e+00
#define A1110 -3.04676447189821950038236690220e+00
#define A1201 2.27331014751653820792359768449e+00
#define A1204 -1.05344954667372501984066689879e+01
#define A1205 -2.00087205822486249909675718444e+00
#define A1206 -1.79589318631187989172765950534e+01
#define A1207 2.79488845294199600508499808837e+01
#define A1208 -2.85899827713502369474065508674e+00
#define A1209 -8.87285693353062954433549289258e+00
#define A1210 1.23605671757943030647266201528e+01
#define A1211 6.43392746015763530355970484046e-01
#define DB01 5.42937341165687622380535766363e-02
#define DB06 4.45031289275240888144113950566e+00
#define DB07 1.89151789931450038304281599044e+00
#define DB08 -5.80120396001058478146721142270e+00
#define DB09 3.11164366957819894408916062370e-01
#define DB10 -1.52160949662516078556178806805e-01
#define DB11 2.01365400804030348374776537501e-01
#define DB12 4.47106157277725905176885569043e-02
#define BH1 0.244094488188976377952755905512e+00
#define BH2 0.733846688281611857341361741547e+00
#define BH3 0.220588235294117647058823529412e-01
#define ER01 0.1312004499419488073250102996e-01
#define ER06 -0.1225156446376204440720569753e+01
#define ER07 -0.4957589496572501915214079952e+00
#define ER08 0.1664377182454986536961530415e+01
#define ER09 -0.3503288487499736816886487290e+00
#define ER10 0.3341791187130174790297318841e+00
#define ER11 0.8192320648511571246570742613e-01
#define ER12 -0.2235530786388629525884427845e-01

// Stage nodes c_i of DOP853 (public dop853.f; the row sums of the stage
// matrix above), for time-dependent fields only: stage i at t + c_i dt.
#define DN02 0.526001519587677318785587544488e-01
#define DN03 0.789002279381515978178381316732e-01
#define DN04 0.118350341907227396726757197510e+00
#define DN05 0.281649658092772603273242802490e+00
#define DN06 0.333333333333333333333333333333e+00
#define DN07 0.25e+00
#define DN08 0.307692307692307692307692307692e+00
#define DN09 0.651282051282051282051282051282e+00
#define DN10 0.6e+00
#define DN11 0.857142857142857142857142857142e+00

// k1 = f(si) is given (raw derivatives, no dt inside, as in the reference);
// returns err and sf.  Slopes are stored in the RHS's type K; stage points,
// sums and the error norm are FP64.
template <class Rhs>
__device__ __forceinline__ double attempt_dp853(Rhs &rhsf, const StepParams &P, const double *si,
                                                const typename Rhs::K *k1, double t, double dt,
                                                double *sf)
{
    using K = typename Rhs::K;
    K k2[6], k3[6], k4[6], k5[6], k6[6], k7[6], k8[6], k9[6], k10[6], k11[6], k12[6];
    double ss[6];
#define RHS(out, node) rhsf.at(t + (node) * dt), rhsf(ss, out)
#pragma unroll
    for (int i = 0; i < 6; i++) ss[i] = si[i] + dt * A0201 * double(k1[i]);
    RHS(k2, DN02);
#pragma unroll
    for (int i = 0; i < 6; i++) ss[i] = si[i] + dt * (A0301 * double(k1[i]) + A0302 * double(k2[i]));
    RHS(k3, DN03);
#pragma unroll
    for (int i = 0; i < 6; i++) ss[i] = si[i] + dt * (A0401 * double(k1[i]) + A0403 * double(k3[i]));
    RHS(k4, DN04);
#pragma unroll
    for (int i = 0; i < 6; i++) ss[i] = si[i] + dt * (A0501 * double(k1[i]) + A0503 * double(k3[i]) + A0504 * double(k4[i]));
    RHS(k5, DN05);
#pragma unroll
    for (int i = 0; i < 6; i++) ss[i] = si[i] + dt * (A0601 * double(k1[i]) + A0604 * double(k4[i]) + A0605 * double(k5[i]));
    RHS(k6, DN06);
#pragma unroll
    for (int i = 0; i < 6; i++)
        ss[i] = si[i] + dt * (A0701 * double(k1[i]) + A0704 * double(k4[i]) + A0705 * double(k5[i]) + A0706 * double(k6[i]));
    RHS(k7, DN07);
#pragma unroll
    for (int i = 0; i < 6; i++)
        ss[i] = si[i] + dt * (A0801 * double(k1[i]) + A0804 * double(k4[i]) + A0805 * double(k5[i]) + A0806 * double(k6[i]) +
                              A0807 * double(k7[i]));
    RHS(k8, DN08);
#pragma unroll
    for (int i = 0; i < 6; i++)
        ss[i] = si[i] + dt * (A0901 * double(k1[i]) + A0904 * double(k4[i]) + A0905 * double(k5[i]) + A0906 * double(k6[i]) +
                              A0907 * double(k7[i]) + A0908 * double(k8[i]));
    RHS(k9, DN09);
#pragma unroll
    for (int i = 0; i < 6; i++)
        ss[i] = si[i] + dt * (A1001 * double(k1[i]) + A1004 * double(k4[i]) + A1005 * double(k5[i]) + A1006 * double(k6[i]) +
                              A1007 * double(k7[i]) + A1008 * double(k8[i]) + A1009 * double(k9[i]));
    RHS(k10, DN10);
#pragma unroll
    for (int i = 0; i < 6; i++)
        ss[i] = si[i] + dt * (A1101 * double(k1[i]) + A1104 * double(k4[i]) + A1105 * double(k5[i]) + A1106 * double(k6[i]) +
                              A1107 * double(k7[i]) + A1108 * double(k8[i]) + A1109 * double(k9[i]) + A1110 * double(k10[i]));
    RHS(k11, DN11);
#pragma unroll
    for (int i = 0; i < 6; i++)
        ss[i] = si[i] + dt * (A1201 * double(k1[i]) + A1204 * double(k4[i]) + A1205 * double(k5[i]) + A1206 * double(k6[i]) +
                              A1207 * double(k7[i]) + A1208 * double(k8[i]) + A1209 * double(k9[i]) + A1210 * double(k10[i]) +
                              A1211 * double(k11[i]));
    rhsf.at(t + dt);
    rhsf(ss, k12);
#undef RHS

    double err3 = 0.0, err5 = 0.0;
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const double kb = DB01 * double(k1[i]) + DB06 * double(k6[i]) + DB07 * double(k7[i]) + DB08 * double(k8[i]) + DB09 * double(k9[i]) +
                          DB10 * double(k10[i]) + DB11 * double(k11[i]) + DB12 * double(k12[i]);
        sf[i] = si[i] + dt * kb;
        const double a = fabs(si[i]), b = fabs(sf[i]);
        const double sr = P.atol + P.rtol * (a > b ? a : b);
        const double e3 = (kb - BH1 * double(k1[i]) - BH2 * double(k9[i]) - BH3 * double(k12[i])) / sr;
        err3 = err3 + e3 * e3;
    }
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const double a = fabs(si[i]), b = fabs(sf[i]);
        const double sr = P.atol + P.rtol * (a > b ? a : b);
        const double e5 = (ER01 * double(k1[i]) + ER06 * double(k6[i]) + ER07 * double(k7[i]) + ER08 * double(k8[i]) + ER09 * double(k9[i]) +
                           ER10 * double(k10[i]) + ER11 * double(k11[i]) + ER12 * double(k12[i])) / sr;
        err5 = err5 + e5 * e5;
    }
    const double deno = err5 + 1.0e-02 * err3;
    return deno <= 0.0 ? fabs(dt) * err5 / sqrt(6.0) : fabs(dt) * err5 / sqrt(6.0 * deno);
}

// cubic Hermite snapshot (integrations.F90:158-168 / :447-457); ka, kb are
// the end slopes times dt
__device__ __forceinline__ void hermite_out(double *st, double t, double tn, double dt,
                                            const double *si, const double *sf, const double *ka,
                                            const double *kb)
{
    const double fr = (t - tn) / dt, fl = 1.0 - fr;
    double o[6];
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const double ds = sf[i] - si[i];
        o[i] = fr * si[i] + fl * sf[i] + fl * fr * ((ka[i] - ds) * fr + (ds - kb[i]) * fl);
    }
#pragma unroll
    for (int i = 0; i < 6; i++) st[i] = o[i];
    st[6] = sqrt(o[3] * o[3] + o[4] * o[4] + o[5] * o[5] + 1.0);
    st[7] = dt;
}

__device__ __forceinline__ bool finite6(const double *a)
{
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 6; i++) ok = ok && isfinite(a[i]);
    return ok;
}

// ---------------------------------------------------------------------------
// persistent work-queue integrator
// ---------------------------------------------------------------------------
// Controller bookkeeping of an accepted step of size h with error ratio err
// (q: order of the error estimate, err ~ h^q).  The record mode completes a
// pending record: its target is this step relative to the one the features
// were taken after.
template <int CTRL>
__device__ __forceinline__ void on_accept(const DeviceParticles &S, int64_t idx, int32_t &flags,
                                          int64_t c_rej, double h, double err, float q,
                                          bool &rej_since, bool &want_feat, double &h_acc,
                                          float &err_acc, bool &rej_before)
{
    if constexpr (CTRL == CTRL_RECORD) {
        if (flags & FLAG_PENDING) {
            const unsigned long long slot = atomicAdd(S.rec_count, 1ull);
            if (slot < (unsigned long long)S.rec_cap) {
                float *__restrict__ r = S.rec + (size_t)kRecord * slot;
                const float *__restrict__ pf = S.pend + (size_t)kFeatures * idx;
#pragma unroll
                for (int i = 0; i < kFeatures; i++) r[i] = pf[i];
                r[16] = float(log(h / S.pend_h[idx]));
                const float le = logf(float(err));
                r[17] = le > -23.0f ? le : -23.0f;
                r[18] = float(c_rej - S.pend_rej[idx]);
                r[19] = q;
            }
            flags &= ~FLAG_PENDING;
        }
    }
    want_feat = true;
    h_acc = h;
    err_acc = float(err);
    rej_before = rej_since;
    rej_since = false;
}

// dynamic LDS of CTRL_LEARNED: the network, then one buffer per wave
constexpr size_t kNetLds = (sizeof(MlpInfer) + 255) & ~size_t(255);

// sampling decision of the record mode for accepted step c of particle idx
__device__ __forceinline__ bool record_sampled(int64_t idx, int64_t c, uint64_t thresh)
{
    uint64_t z = (uint64_t)idx * 0x9e3779b97f4a7c15ull + (uint64_t)c;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return (z >> 32) < thresh;
}

// CTRL selects the step-size controller (CTRL_PI: the reference's, bit for
// bit; CTRL_RECORD: the same, writing sampled training records; CTRL_LEARNED:
// after an accepted step the network proposes the next step, after a
// rejection the PI rule does).  CTRL_LEARNED takes the network and a 64x72
// BF16 buffer per wave in dynamic LDS.
// Cell-cache values the static FP64 PI kernels keep in LDS (per-wave slabs,
// 64 doubles per value) instead of registers.  DOP853 is at the 512-register
// limit: with all 48 values in registers it spills 212 B per lane; with the
// last 16 in LDS it spills 36 B and runs 1.09x faster on the 256^3
// turbulent workload (8 to 20 measured alike, 32 or more far slower: the
// sums then wait on too many LDS reads).  RK5 spills 52 B and gains nothing.
// The FP32 kernel does not spill (its overflow goes to AGPRs); moving 16 or
// 32 of its values to LDS made it 4 % and 40 % slower.
constexpr int static_split(int method, int ctrl, int cache, int prec, int time)
{
    return method == PTC_DOP853 && ctrl == CTRL_PI && cache == 1 && prec == PTC_FP64 && time == 0
               ? 16
               : 0;
}

// TIME = 1 integrates in a snapshot series (fields interpolated in time at
// each stage; PI controller, node rows, cell cache as described at Force).
// PARK = 1 (static node rows, PI controller): once the queue is empty
// and at most S.park particles are still in flight (S.qcount[2] counts the
// finished ones and those fetched already retired), every wave stores its particles as at the end of a segment
// and stops; the group kernel (group_impl.hpp) finishes them.  A separate
// instantiation, because the extra live state changes the register
// allocation of the kernels that do not hand over.
template <int METHOD, int MODE, int LAYOUT, int CACHE, int PREC, int CTRL, int GEOM, int TIME = 0,
          int PARK = 0>
__global__ void __launch_bounds__(kBlock)
integrate_kernel(const DeviceFields F, const StepParams P, const DeviceParticles S, double t_stop)
{
    static_assert(PARK == 0 || (CTRL == CTRL_PI && TIME == 0 && LAYOUT == LAYOUT_AOS),
                  "the group kernel takes static node rows with the PI controller");
    static_assert(CTRL == CTRL_PI || (CACHE == 1 && LAYOUT == LAYOUT_AOS),
                  "step-size features read the cell cache");
    static_assert(TIME == 0 || (CTRL == CTRL_PI && LAYOUT == LAYOUT_AOS),
                  "a snapshot series runs with the PI controller on node rows");
    // the power's lookup tables (5 KB) in LDS: per-lane indices make every
    // lookup a divergent gather, which LDS serves far faster than memory
    __shared__ double s_logt[128][3];
    __shared__ unsigned long long s_expt[256];
    for (int i = threadIdx.x; i < 128 * 3; i += blockDim.x)
        (&s_logt[0][0])[i] = (&ptc_pow_log_tab[0][0])[i];
    for (int i = threadIdx.x; i < 256; i += blockDim.x) s_expt[i] = ptc_pow_exp_tab[i];
    extern __shared__ __attribute__((aligned(16))) char s_dyn[];
    if constexpr (CTRL == CTRL_LEARNED) {
        for (int i = threadIdx.x; i < (int)(sizeof(MlpInfer) / 4); i += blockDim.x)
            ((uint32_t *)s_dyn)[i] = ((const uint32_t *)&S.net->f)[i];
    }
    __syncthreads();
    const PowTables T{s_logt, s_expt};

    constexpr int kSplit = static_split(METHOD, CTRL, CACHE, PREC, TIME);
    Force<MODE, LAYOUT, CACHE, PREC, GEOM, TIME, kSplit> force(F, P);
    using K = typename decltype(force)::K;
    if constexpr (TIME == 1 && CACHE == 1)
        force.lds1 = (K *)s_dyn + (threadIdx.x >> 6) * (48 * 64) + (threadIdx.x & 63);
    if constexpr (kSplit > 0)
        force.lds1 = (K *)s_dyn + (threadIdx.x >> 6) * (kSplit * 64) + (threadIdx.x & 63);
    const int lane = (int)__lane_id();
    const double t_end = t_stop < P.tmx ? t_stop : P.tmx;
    const int64_t N = S.np;

    bool active = false, exhausted = false, need_rhs = false, fresh = false;
    int64_t idx = 0;
    double si[6];
    K f1[6];
    double t = 0.0, dt = 0.0, errold = 1.0e-04, tnext = 0.0;
    int32_t n = 2, flags = 0, status = PTC_STATUS_RUNNING;
    int64_t c_all = 0, c_acc = 0, c_rej = 0, c_rhs = 0;
    // controller state: features are taken at the top of the iteration after
    // an accepted step of size h_acc with error ratio err_acc
    bool want_feat = false, rej_since = false, rej_before = false;
    double h_acc = 0.0;
    float err_acc = 0.0f, pi_ratio = 0.0f;

    for (;;) {
        // ---- refill lanes without a particle (one atomic per wave) ----
        const bool want = !active && !exhausted;
        const uint64_t wmask = __ballot(want);
        bool retired = false;  // fetched a particle that no longer runs
        if (wmask) {
            const int leader = __ffsll((unsigned long long)wmask) - 1;
            int base = 0;
            if (lane == leader) base = atomicAdd(S.qcount, __popcll(wmask));
            base = __shfl(base, leader);
            if (want) {
                const int64_t my = (int64_t)base + __popcll(wmask & ((1ull << lane) - 1ull));
                if (my >= N) {
                    exhausted = true;
                } else {
                    idx = S.order ? (int64_t)S.order[my] : my;
                    status = S.status[idx];
                    retired = status != PTC_STATUS_RUNNING;
                    if (status == PTC_STATUS_RUNNING) {
                        active = true;
                        fresh = true;
                        need_rhs = true;
                        want_feat = rej_since = false;
#pragma unroll
                        for (int c = 0; c < 6; c++) si[c] = S.x[c][idx];
                        t = S.t[idx];
                        dt = S.dt[idx];
                        errold = S.errold[idx];
                        n = S.nout[idx];
                        flags = S.flags[idx];
                        c_all = S.cnt[idx];
                        c_acc = S.cnt[N + idx];
                        c_rej = S.cnt[2 * N + idx];
                        c_rhs = S.nrhs[idx];
                    }
                }
            }
        }
        if constexpr (PARK == 1) {
            // particles retired before this launch (done, capped, non-finite)
            // count as finished, so the hand-over threshold compares the
            // particles still in flight
            const uint64_t rm = __ballot(retired);
            if (rm && lane == __ffsll((unsigned long long)rm) - 1)
                atomicAdd(S.qcount + 2, __popcll(rm));
            if (__any(exhausted)) {
                int done = 0;
                if (lane == 0)
                    done = __hip_atomic_load(S.qcount + 2, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
                done = __shfl(done, 0);
                if (N - done <= (int64_t)S.park) {
                    if (active) {
#pragma unroll
                        for (int c = 0; c < 6; c++) S.x[c][idx] = si[c];
                        S.t[idx] = t;
                        S.dt[idx] = dt;
                        S.errold[idx] = errold;
                        S.nout[idx] = n;
                        S.flags[idx] = flags;
                        S.cnt[idx] = c_all;
                        S.cnt[N + idx] = c_acc;
                        S.cnt[2 * N + idx] = c_rej;
                        S.nrhs[idx] = c_rhs;
                    }
                    break;
                }
            }
        }
        if (!__any(active)) {
            if (__all(exhausted)) break;
            continue;
        }

        // ---- one shared RHS slot: f(si) for refilled lanes and for lanes
        //      that accepted last iteration (the next step's stage 1) ----
        // (step-size modes: the same evaluation also yields the features of
        // the accepted state, from the same cell and interpolated fields)
        float feat[CTRL == CTRL_PI ? 1 : kFeatures];
        bool feat_ready = false;
        const bool feat_wanted =
            CTRL == CTRL_LEARNED ? want_feat
                                 : want_feat && record_sampled(idx, c_acc, S.rec_thresh);
        if (need_rhs) {
            force.at(t);
            // (FP64 fields: the kernel is short of registers already, the
            // features are taken from the cache afterwards instead)
            if constexpr (CTRL == CTRL_PI || PREC != PTC_FP32) {
                force(si, f1);
            } else {
                float ub[6], xd[3];
                force.eval_fields(si, f1, ub, xd);
                if (feat_wanted) {
                    features_core<MODE>(P, si, ub, force.v, xd, h_acc, err_acc, pi_ratio,
                                        rej_before, feat);
                    feat_ready = true;
                }
            }
            c_rhs++;
            need_rhs = false;
        }

        if constexpr (CTRL == CTRL_RECORD) {
            if (feat_wanted) {
                if (!feat_ready)
                    step_features<MODE>(force, P, si, h_acc, err_acc, pi_ratio, rej_before, feat);
                float *__restrict__ pf = S.pend + (size_t)kFeatures * idx;
#pragma unroll
                for (int i = 0; i < kFeatures; i++) pf[i] = feat[i];
                S.pend_h[idx] = h_acc;
                S.pend_rej[idx] = c_rej;
                flags |= FLAG_PENDING;
            }
            want_feat = false;
        } else if constexpr (CTRL == CTRL_LEARNED) {
            if (__any(want_feat)) {
                if (want_feat) {
                    if (!feat_ready)
                        step_features<MODE>(force, P, si, h_acc, err_acc, pi_ratio, rej_before,
                                            feat);
                } else {
#pragma unroll
                    for (int i = 0; i < kFeatures; i++) feat[i] = 0.0f;
                }
                const MlpInfer &W = *(const MlpInfer *)s_dyn;
                MlpWaveBuf &wb = ((MlpWaveBuf *)(s_dyn + kNetLds))[threadIdx.x >> 6];
                const float y = mlp_forward_wave(W, wb, feat);
                if (want_feat && isfinite(y)) {
                    // growth at most facmax; never below facmin times the PI
                    // proposal, so a poor network cannot stall a particle
                    float lg = y + S.ctrl_shift;
                    const float lfmn = logf(float(P.fcmn)), hi = logf(float(P.fcmx));
                    const float lo = pi_ratio + lfmn > lfmn ? pi_ratio + lfmn : lfmn;
                    lg = lg < lo ? lo : lg;
                    lg = lg > hi ? hi : lg;
                    const double dtl = h_acc * double(__expf(lg));
                    dt = dtl < P.dtm ? dtl : P.dtm;
                }
            }
            want_feat = false;
        }

        bool finish = false;
        if (fresh) {
            fresh = false;
            if (!(flags & FLAG_STARTED)) {
                // integrations.F90:120-127 / :352-359
                dt = dti_guess(P, METHOD == PTC_RK5 ? 5 : 8, si, f1, T);
                double *o = S.out + (size_t)8 * P.m * idx;
                o[6] = sqrt(si[3] * si[3] + si[4] * si[4] + si[5] * si[5] + 1.0);
                o[7] = dt;
                t = 0.0;
                n = 2;
                errold = 1.0e-04;
                flags = FLAG_STARTED;
                c_all = c_acc = c_rej = 0;
            }
            tnext = n <= P.m ? S.time[n - 1] : 0.0;
            if (!(t < t_end)) {
                if (!(t < P.tmx)) status = PTC_STATUS_DONE;
                finish = true;
            }
        }

        if (active && !finish) {
            if (P.max_steps > 0 && c_all >= P.max_steps) {
                status = PTC_STATUS_MAXSTEPS;
                finish = true;
            } else if (!isfinite(dt)) {
                status = PTC_STATUS_NONFINITE;
                finish = true;
            } else if (!(dt > 0.0) || t + dt == t) {
                status = PTC_STATUS_DTSMALL;
                finish = true;
            } else {
                double sf[6];
                double dtn;
                if constexpr (METHOD == PTC_RK5) {
                    double k6[6];
                    const double del = attempt_rk5(force, P, si, f1, t, dt, sf, k6);
                    c_rhs += 5;
                    if (!isfinite(del) || !finite6(sf)) {
                        status = PTC_STATUS_NONFINITE;
                        finish = true;
                    } else {
                      // one power per lane serves both outcomes:
                      // del**(-0.2) on accept, del**(-0.25) on reject
                      // (integrations.F90:177,186)
                      const bool accept = del <= 1.0;
                      const double q = P.safe * ptc_pow_with(del, accept ? -0.20 : -0.25, T.log, T.exp);
                      if (accept) {
                        t = t + dt;
                        if (n <= P.m && t >= tnext) {
                            double k1[6];
#pragma unroll
                            for (int c = 0; c < 6; c++) k1[c] = dt * double(f1[c]);
                            do {
                                hermite_out(S.out + (size_t)8 * ((n - 1) + (size_t)P.m * idx), t,
                                            tnext, dt, si, sf, k1, k6);
                                n++;
                                tnext = n <= P.m ? S.time[n - 1] : 0.0;
                            } while (n <= P.m && t >= tnext);
                        }
#pragma unroll
                        for (int c = 0; c < 6; c++) si[c] = sf[c];
                        if constexpr (CTRL != CTRL_PI)
                            on_accept<CTRL>(S, idx, flags, c_rej, dt, del, 5.0f, rej_since,
                                            want_feat, h_acc, err_acc, rej_before);
                        c_acc++;
                        dtn = dt * (P.fcmx < q ? P.fcmx : q);
                        need_rhs = true;
                      } else {
                        c_rej++;
                        rej_since = true;
                        dtn = dt * (P.fcmn > q ? P.fcmn : q);
                      }
                    }
                } else {
                    const double err = attempt_dp853(force, P, si, f1, t, dt, sf);
                    c_rhs += 11;
                    if (!isfinite(err) || !finite6(sf)) {
                        status = PTC_STATUS_NONFINITE;
                        finish = true;
                    } else {
                      // safe * err**expo serves both outcomes (integrations.F90:466,486)
                      const double qe = P.safe * ptc_pow_with(err, P.expo, T.log, T.exp);
                      if (err <= 1.0) {
                        t = t + dt;
                        if (n <= P.m && t >= tnext) {
                            // the end slope f(sf) is needed now; it is also
                            // the next step's stage 1
                            K fe[6];
                            double ka[6], kb[6];
                            force.at(t);
                            force(sf, fe);
                            c_rhs++;
#pragma unroll
                            for (int c = 0; c < 6; c++) {
                                ka[c] = dt * double(f1[c]);
                                kb[c] = dt * double(fe[c]);
                            }
                            do {
                                hermite_out(S.out + (size_t)8 * ((n - 1) + (size_t)P.m * idx), t,
                                            tnext, dt, si, sf, ka, kb);
                                n++;
                                tnext = n <= P.m ? S.time[n - 1] : 0.0;
                            } while (n <= P.m && t >= tnext);
#pragma unroll
                            for (int c = 0; c < 6; c++) f1[c] = fe[c];
                        } else {
                            need_rhs = true;
                        }
#pragma unroll
                        for (int c = 0; c < 6; c++) si[c] = sf[c];
                        if constexpr (CTRL != CTRL_PI)
                            on_accept<CTRL>(S, idx, flags, c_rej, dt, err, float(-1.0 / P.expo),
                                            rej_since, want_feat, h_acc, err_acc, rej_before);
                        c_acc++;
                        double q = qe;
                        // errold**beta; x**0 is exactly 1 (and q * 1 == q),
                        // so the beta = 0 default skips the power
                        if (P.beta != 0.0) q = q * ptc_pow_with(errold, P.beta, T.log, T.exp);
                        if (q < P.fcmn) q = P.fcmn;
                        if (flags & FLAG_REJECTED) dtn = dt * (q < 1.0 ? q : 1.0);
                        else dtn = dt * (q < P.fcmx ? q : P.fcmx);
                        errold = err > 1.0e-04 ? err : 1.0e-04;
                        flags &= ~FLAG_REJECTED;
                      } else {
                        c_rej++;
                        rej_since = true;
                        dtn = dt * (P.fcmn > qe ? P.fcmn : qe);
                        flags |= FLAG_REJECTED;
                      }
                    }
                }
                if (!finish) {
                    c_all++;
                    dt = dtn < P.dtm ? dtn : P.dtm;
                    if constexpr (CTRL != CTRL_PI) {
                        if (want_feat) pi_ratio = float(log(dt / h_acc));
                    }
                    if (!(t < t_end)) {
                        if (!(t < P.tmx)) status = PTC_STATUS_DONE;
                        finish = true;
                    }
                }
            }
        }

        if constexpr (PARK == 1) {
            const uint64_t fm = __ballot(active && finish);
            if (fm && lane == __ffsll((unsigned long long)fm) - 1)
                atomicAdd(S.qcount + 2, __popcll(fm));
        }
        if (active && finish) {
#pragma unroll
            for (int c = 0; c < 6; c++) S.x[c][idx] = si[c];
            S.t[idx] = t;
            S.dt[idx] = dt;
            S.errold[idx] = errold;
            S.nout[idx] = n;
            S.flags[idx] = flags;
            S.status[idx] = status;
            S.cnt[idx] = c_all;
            S.cnt[N + idx] = c_acc;
            S.cnt[2 * N + idx] = c_rej;
            S.nrhs[idx] = c_rhs;
            active = false;
            need_rhs = false;
            want_feat = false;
        }
    }
    add_cell_loads(S, force.loads);
}

// Single-call probes of the device building blocks, one input per thread:
//   op 0: gather at pos(3)          -> f(6)  (U or E, then B)
//   op 1: RHS at vec(6)             -> acc(6)
//   op 2: dti_guess, order 5, vec(6) -> dt(1) (op 3: order 8)
//   op 4: step-size power x^y, (x, y) -> (1)
//   op 5: step-size features, (vec(6), h, err, pi_ratio, rejected) -> (16)
//   op 6: gather at (pos(3), t) in a snapshot series -> f(6)   (node rows)
//   op 7: RHS at (vec(6), t) in a snapshot series    -> acc(6) (node rows)
template <int MODE, int LAYOUT, int PREC>
__global__ void probe_kernel(const DeviceFields F, const StepParams P, int op, int64_t n,
                             const double *__restrict__ in, double *__restrict__ out)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v[6];
    Force<MODE, LAYOUT, 0, PREC> force(F, P);
    typename decltype(force)::K r[6];
    if (op == 0) {
#pragma unroll
        for (int k = 0; k < 3; k++) v[k] = in[3 * i + k];
        if constexpr (PREC == PTC_FP32) {
            CornersT<float> c;
            corners(P, v, c);
            float f[6];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                float w[6];
                load_node_f32((const float *)F.data, c.o[j], w);
                accumulate_row(j, c.w[j], w, f);
            }
#pragma unroll
            for (int k = 0; k < 6; k++) r[k] = f[k];
        } else {
            gather<LAYOUT>(F, P, v, r);
        }
#pragma unroll
        for (int k = 0; k < 6; k++) out[6 * i + k] = double(r[k]);
    } else if (op == 1) {
#pragma unroll
        for (int k = 0; k < 6; k++) v[k] = in[6 * i + k];
        force(v, r);
#pragma unroll
        for (int k = 0; k < 6; k++) out[6 * i + k] = double(r[k]);
    } else if (op == 2 || op == 3) {
#pragma unroll
        for (int k = 0; k < 6; k++) v[k] = in[6 * i + k];
        force(v, r);
        out[i] = dti_guess(P, op == 2 ? 5 : 8, v, r, PowTables{ptc_pow_log_tab, ptc_pow_exp_tab});
    } else if (op == 4) {
        out[i] = ptc_libm_pow(in[2 * i], in[2 * i + 1]);
    } else if (op == 5) {
        if constexpr (LAYOUT == LAYOUT_AOS) {
            const double *q = in + 10 * i;
#pragma unroll
            for (int k = 0; k < 6; k++) v[k] = q[k];
            Force<MODE, LAYOUT_AOS, 1, PREC> fc(F, P);
            float feat[kFeatures];
            step_features<MODE>(fc, P, v, q[6], float(q[7]), float(q[8]), q[9] != 0.0, feat);
#pragma unroll
            for (int k = 0; k < kFeatures; k++) out[kFeatures * i + k] = double(feat[k]);
        }
    } else if (op == 6 || op == 7) {
        if constexpr (LAYOUT == LAYOUT_AOS) {
            Force<MODE, LAYOUT_AOS, 0, PREC, 0, 1> ft(F, P);
            const int w = op == 6 ? 3 : 6;
            const double *q = in + (size_t)(w + 1) * i;
            for (int k = 0; k < 6; k++) v[k] = k < w ? q[k] : 0.0;
            ft.at(q[w]);
            if (op == 6) ft.fields_at(v, r);
            else ft(v, r);
#pragma unroll
            for (int k = 0; k < 6; k++) out[6 * i + k] = double(r[k]);
        }
    }
}

// dynamic LDS of a cached series kernel: snapshot k+1's 48 corner values
// per thread
inline size_t series_cache_lds(int prec, int block)
{
    const size_t waves = ((size_t)block + 63) / 64;
    return (size_t)48 * 64 * (prec == PTC_FP32 ? sizeof(float) : sizeof(double)) * waves;
}

// whether the device gives one block that much dynamic LDS (next to the
// kernel's 5 KB of static tables)
inline bool series_cache_fits(int prec, int block)
{
    int dev = 0, max_lds = 0;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
        hipSuccess)
        return false;
    return series_cache_lds(prec, block) + 8192 <= (size_t)max_lds;
}

template <int MODE, int LAYOUT, int PREC = PTC_FP64>
hipError_t launch_probe_t(const DeviceFields &f, const StepParams &sp, int op, int64_t n,
                          const double *in, double *out, hipStream_t stream)
{
    const int64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL((probe_kernel<MODE, LAYOUT, PREC>), dim3((unsigned)blocks), dim3(256), 0,
                       stream, f, sp, op, n, in, out);
    return hipGetLastError();
}

template <int METHOD, int MODE, int LAYOUT, int CACHE, int PREC, int CTRL, int GEOM, int TIME = 0,
          int PARK = 0>
hipError_t launch_t(const DeviceFields &f, const StepParams &sp, const DeviceParticles &p,
                    double t_stop, LaunchConfig lc, hipStream_t stream)
{
    auto kern = integrate_kernel<METHOD, MODE, LAYOUT, CACHE, PREC, CTRL, GEOM, TIME, PARK>;
    constexpr int kSplit = static_split(METHOD, CTRL, CACHE, PREC, TIME);
    const size_t lds = CTRL == CTRL_LEARNED ? kNetLds + (size_t)(lc.block / 64) * sizeof(MlpWaveBuf)
                       : kSplit > 0 ? (size_t)kSplit * 64 * sizeof(double) * (size_t)((lc.block + 63) / 64)
                       : TIME == 1 && CACHE == 1 && PREC == PTC_FP64
                           ? series_cache_lds(PREC, lc.block)
                           : 0;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void *)kern,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    if (CTRL == CTRL_LEARNED && !p.net) return hipErrorInvalidValue;
    int grid = lc.grid_blocks;
    if (grid <= 0) {
        int dev = 0, cus = 0, per_cu = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e != hipSuccess) return e;
        e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return e;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, lc.block, lds);
        if (e != hipSuccess) return e;
        grid = cus * (per_cu > 0 ? per_cu : 1);
    }
    const int64_t need = (p.np + lc.block - 1) / lc.block;
    if ((int64_t)grid > need) grid = (int)need;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(lc.block), lds, stream, f, sp, p, t_stop);
    return hipGetLastError();
}

}  // namespace
}  // namespace ptc
